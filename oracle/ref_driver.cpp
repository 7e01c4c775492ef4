// ref_driver.cpp -- test-only harness that links ls1-MarDyn's own molecule
// sources (FullMolecule.cpp, Component.cpp, Comp2Param.cpp, Quaternion.cpp,
// mixingrules, Logger, xmlfile*) and its scalar pair physics (potforce.h, the
// "Legacy" path that VectorizedCellProcessorTest compares VCP against,
// src/particleContainer/adapter/tests/VectorizedCellProcessorTest.cpp:136-219).
// Built by oracle/Makefile into oracle/_ref/libls1ref.so from the sources
// where they lie; nothing of the reference is copied into this repository.
// TEST INFRASTRUCTURE ONLY: used to pin the C restatement in ls1_oracle.c.
#include <cmath>
#include <memory>
#include <vector>

#include "molecules/Comp2Param.h"
#include "molecules/Component.h"
#include "molecules/Molecule.h"
#include "molecules/mixingrules/LorentzBerthelot.h"
#include "molecules/potforce.h"

namespace {
struct RefModel {
    std::vector<Component> comps;
    MixingRuleBase::MixRulesType mix;
    Comp2Param c2p;
    double epsRF = 1.0, rc = 1.0, rcLJ = 1.0;
};

// Mirrors the ASCII header reading sequence (src/io/ASCIIReader.cpp:152-232).
int parse(RefModel &m, int ncomp, const double *desc, int len, const double *mixc) {
    int p = 0;
    m.comps.resize(ncomp);
    for (int i = 0; i < ncomp; ++i) {
        Component &c = m.comps[i];
        c.setID(i);
        int nlj = (int)desc[p], nc = (int)desc[p + 1], nd = (int)desc[p + 2], nq = (int)desc[p + 3];
        p += 4;
        for (int s = 0; s < nlj; ++s, p += 8)
            c.addLJcenter(desc[p], desc[p + 1], desc[p + 2], desc[p + 3], desc[p + 4], desc[p + 5],
                          desc[p + 6], desc[p + 7] != 0.);
        for (int s = 0; s < nc; ++s, p += 5)
            c.addCharge(desc[p], desc[p + 1], desc[p + 2], desc[p + 3], desc[p + 4]);
        for (int s = 0; s < nd; ++s, p += 7)
            c.addDipole(desc[p], desc[p + 1], desc[p + 2], desc[p + 3], desc[p + 4], desc[p + 5],
                        desc[p + 6]);
        for (int s = 0; s < nq; ++s, p += 7)
            c.addQuadrupole(desc[p], desc[p + 1], desc[p + 2], desc[p + 3], desc[p + 4],
                            desc[p + 5], desc[p + 6]);
        if (desc[p] > 0.) c.setI11(desc[p]);
        if (desc[p + 1] > 0.) c.setI22(desc[p + 1]);
        if (desc[p + 2] > 0.) c.setI33(desc[p + 2]);
        p += 3;
        if (p > len) return -1;
    }
    int k = 0;
    for (int i = 0; i < ncomp; ++i)
        for (int j = i + 1; j < ncomp; ++j, ++k) {
            auto rule = std::make_shared<LorentzBerthelotMixingRule>();
            rule->setCid1(i);
            rule->setCid2(j);
            rule->setXi(mixc ? mixc[2 * k] : 1.0);
            rule->setEta(mixc ? mixc[2 * k + 1] : 1.0);
            m.mix[i][j] = rule;
        }
    return 0;
}
}  // namespace

extern "C" {

void *ref_create(int ncomp, const double *desc, int len, const double *mix, double epsRF,
                 double rc, double rcLJ) {
    RefModel *m = new RefModel();
    if (parse(*m, ncomp, desc, len, mix)) {
        delete m;
        return nullptr;
    }
    m->epsRF = epsRF;
    m->rc = rc;
    m->rcLJ = rcLJ;
    m->c2p.initialize(m->comps, m->mix, epsRF, rc, rcLJ);
    return m;
}

void ref_destroy(void *h) { delete static_cast<RefModel *>(h); }

void ref_component_props(void *h, double *mass, double *inertia, int *rotdof) {
    RefModel *m = static_cast<RefModel *>(h);
    for (size_t i = 0; i < m->comps.size(); ++i) {
        mass[i] = m->comps[i].m();
        inertia[3 * i] = m->comps[i].I11();
        inertia[3 * i + 1] = m->comps[i].I22();
        inertia[3 * i + 2] = m->comps[i].I33();
        rotdof[i] = (int)m->comps[i].getRotationalDegreesOfFreedom();
    }
}

// The VCP parameter table layout (row/column = LJ lookup id), read from the
// reference's Comp2Param streams in the order its consumers read them.
void ref_lj_tables(void *h, double *eps24, double *sig2, double *shift6) {
    RefModel *m = static_cast<RefModel *>(h);
    int n = 0;
    std::vector<int> base;
    for (auto &c : m->comps) {
        base.push_back(n);
        n += c.numLJcenters();
    }
    for (size_t ci = 0; ci < m->comps.size(); ++ci)
        for (size_t cj = 0; cj < m->comps.size(); ++cj) {
            ParaStrm &p = m->c2p(ci, cj);
            p.reset_read();
            for (unsigned a = 0; a < m->comps[ci].numLJcenters(); ++a)
                for (unsigned b = 0; b < m->comps[cj].numLJcenters(); ++b) {
                    double e, s, sh;
                    p >> e;
                    p >> s;
                    p >> sh;
                    int k = (base[ci] + a) * n + (base[cj] + b);
                    eps24[k] = e;
                    sig2[k] = s;
                    shift6[k] = sh;
                }
        }
}

// Legacy-path forces: every molecule pair with centre distance below rc
// (LegacyCellProcessor.cpp:53-154 criterion), evaluated by potforce.h's
// PotForce with the ParticlePairs2PotForceAdapter accounting
// (ParticlePairs2PotForceAdapter.h:151-160), then FullMolecule::calcFM.
// periodic != 0 uses the minimum image (box > 2 rc). macro = upot, virial.
int ref_forces(void *h, int n, const double *r, const double *q, const int *cid,
               const double *box, int periodic, double *F, double *M, double *Vi, double *macro) {
    RefModel *m = static_cast<RefModel *>(h);
    std::vector<Molecule> mols;
    mols.reserve(n);
    for (int i = 0; i < n; ++i) {
        mols.emplace_back(i, &m->comps[cid[i]], r[3 * i], r[3 * i + 1], r[3 * i + 2], 0., 0., 0.,
                          q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3], 0., 0., 0.);
    }
    for (auto &mo : mols) mo.buildOwnSoA();
    std::vector<double> extra(9 * (size_t)n, 0.0);  // F, M, Vi from periodic images
    double upot6 = 0, upotX = 0, myRF = 0, virial = 0;
    const double rc2 = m->rc * m->rc, rcLJ2 = m->rcLJ * m->rcLJ;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double s[3] = {0, 0, 0};
            if (periodic)
                for (int d = 0; d < 3; ++d) {
                    double dd = r[3 * i + d] - r[3 * j + d];
                    if (dd > 0.5 * box[d]) s[d] = box[d];
                    else if (dd < -0.5 * box[d]) s[d] = -box[d];
                }
            bool img = s[0] != 0 || s[1] != 0 || s[2] != 0;
            Molecule *pj = &mols[j];
            Molecule copy;
            if (img) {
                copy = Molecule(j, &m->comps[cid[j]], r[3 * j] + s[0], r[3 * j + 1] + s[1],
                                r[3 * j + 2] + s[2], 0., 0., 0., q[4 * j], q[4 * j + 1],
                                q[4 * j + 2], q[4 * j + 3], 0., 0., 0.);
                copy.buildOwnSoA();
                pj = &copy;
            }
            double drm[3];
            double dd = pj->dist2(mols[i], drm);
            if (!(dd < rc2)) continue;
            ParaStrm &params = m->c2p(cid[i], cid[j]);
            params.reset_read();
            double V3[3];
            PotForce(mols[i], *pj, params, drm, upot6, upotX, myRF, V3, dd < rcLJ2);
            virial += 2 * (V3[0] + V3[1] + V3[2]);
            for (int d = 0; d < 3; ++d) {
                extra[9 * i + 6 + d] += V3[d];
                extra[9 * j + 6 + d] += V3[d];
            }
            if (img) {
                copy.calcFM();
                for (int d = 0; d < 3; ++d) {
                    extra[9 * j + d] += copy.F(d);
                    extra[9 * j + 3 + d] += copy.M(d);
                }
                copy.releaseOwnSoA();
            }
        }
    for (int i = 0; i < n; ++i) {
        mols[i].calcFM();
        for (int d = 0; d < 3; ++d) {
            F[3 * i + d] = mols[i].F(d) + extra[9 * i + d];
            M[3 * i + d] = mols[i].M(d) + extra[9 * i + 3 + d];
            Vi[3 * i + d] = extra[9 * i + 6 + d];
        }
        mols[i].releaseOwnSoA();
    }
    // ParticlePairs2PotForceAdapter::finish, ParticlePairs2PotForceAdapter.h:93-98
    macro[0] = upot6 / 6. + upotX + myRF;
    macro[1] = virial + 3.0 * myRF;
    return 0;
}

// FullMolecule::upd_preF for n molecules (Leapfrog::transition1to2).
void ref_upd_preF(void *h, int n, double dt, double *r, double *v, double *q, double *L,
                  const double *F, const double *M, const int *cid) {
    RefModel *m = static_cast<RefModel *>(h);
    for (int i = 0; i < n; ++i) {
        Molecule mo(i, &m->comps[cid[i]], r[3 * i], r[3 * i + 1], r[3 * i + 2], v[3 * i],
                    v[3 * i + 1], v[3 * i + 2], q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3],
                    L[3 * i], L[3 * i + 1], L[3 * i + 2]);
        double f[3] = {F[3 * i], F[3 * i + 1], F[3 * i + 2]};
        double mm[3] = {M[3 * i], M[3 * i + 1], M[3 * i + 2]};
        mo.setF(f);
        mo.setM(mm);
        mo.upd_preF(dt);
        for (int d = 0; d < 3; ++d) {
            r[3 * i + d] = mo.r(d);
            v[3 * i + d] = mo.v(d);
            L[3 * i + d] = mo.D(d);
        }
        q[4 * i] = mo.q().qw();
        q[4 * i + 1] = mo.q().qx();
        q[4 * i + 2] = mo.q().qy();
        q[4 * i + 3] = mo.q().qz();
    }
}

// FullMolecule::upd_postF for n molecules; sums = (summv2, sumIw2).
void ref_upd_postF(void *h, int n, double dt, double *v, double *L, const double *q,
                   const double *F, const double *M, const int *cid, double *sums) {
    RefModel *m = static_cast<RefModel *>(h);
    double summv2 = 0, sumIw2 = 0;
    for (int i = 0; i < n; ++i) {
        Molecule mo(i, &m->comps[cid[i]], 0., 0., 0., v[3 * i], v[3 * i + 1], v[3 * i + 2],
                    q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3], L[3 * i], L[3 * i + 1],
                    L[3 * i + 2]);
        double f[3] = {F[3 * i], F[3 * i + 1], F[3 * i + 2]};
        double mm[3] = {M[3 * i], M[3 * i + 1], M[3 * i + 2]};
        mo.setF(f);
        mo.setM(mm);
        mo.upd_postF(0.5 * dt, summv2, sumIw2);
        for (int d = 0; d < 3; ++d) {
            v[3 * i + d] = mo.v(d);
            L[3 * i + d] = mo.D(d);
        }
    }
    sums[0] = summv2;
    sums[1] = sumIw2;
}
}
