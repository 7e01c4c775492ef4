// HIPLinkedCells.cpp -- see HIPLinkedCells.h.
#include "HIPLinkedCells.h"

#include <cmath>

namespace ls1amd {

namespace {
[[noreturn]] void die(const std::string &msg) { throw std::runtime_error(msg); }
}  // namespace

void HIPLinkedCells::check(int rc, const char *what) {
    if (rc != MDX_OK) die(std::string(what) + " failed: " + mdx_last_error(_ctx));
}

HIPLinkedCells::HIPLinkedCells(double bBoxMin[3], double bBoxMax[3], double cutoffRadius,
                               double LJcutoffRadius, const ComponentModel &model, bool periodic,
                               int device, int precision)
    : _model(model), _cutoff(cutoffRadius), _ljCutoff(LJcutoffRadius), _periodic(periodic),
      _device(device), _precision(precision) {
    for (int d = 0; d < 3; ++d) {
        _bbMin[d] = bBoxMin[d];
        _bbMax[d] = bBoxMax[d];
    }
    makeContext(0);
}

// A context on the container's device with its model, domain, precision and
// the thermostat / boundary settings made through this class.
void HIPLinkedCells::makeContext(int cellMode) {
    if (_ctx) mdx_destroy(_ctx);
    _ctx = nullptr;
    if (mdx_create(_device, &_ctx) != MDX_OK)
        die(std::string("mdx_create failed: ") + mdx_last_error(nullptr));
    check(mdx_set_components(_ctx, _model.ncomp, _model.desc.data(), (int)_model.desc.size(),
                             _model.mix.empty() ? nullptr : _model.mix.data(), _model.epsilonRF),
          "mdx_set_components");
    setDomain(cellMode);
    check(mdx_set_precision(_ctx, _precision), "mdx_set_precision");
    if (_haveBoundaries) check(mdx_set_boundaries(_ctx, _boundaries), "mdx_set_boundaries");
    if (_haveThermostat)
        check(mdx_set_thermostat(_ctx, 1, &_targetT, _lrc ? 1 : 0, _heuristics ? 1 : 0),
              "mdx_set_thermostat");
}

HIPLinkedCells::~HIPLinkedCells() { mdx_destroy(_ctx); }

void HIPLinkedCells::setDomain(int cellMode) {
    for (int d = 0; d < 3; ++d)
        if (_bbMin[d] != 0.0)
            die("HIPLinkedCells: the single-GPU domain starts at the origin (bBoxMin = 0)");
    double box[3] = {_bbMax[0] - _bbMin[0], _bbMax[1] - _bbMin[1], _bbMax[2] - _bbMin[2]};
    check(mdx_set_domain(_ctx, box, _cutoff, _ljCutoff, _periodic ? 1 : 0, cellMode),
          "mdx_set_domain");
}

// LinkedCells::rebuild (LinkedCells.cpp:132-201) uses a float cell edge. The
// molecules are kept; the new context keeps the device, precision and the
// settings made through this class (settings made directly on context() are
// not carried over).
bool HIPLinkedCells::rebuild(double bBoxMin[3], double bBoxMax[3]) {
    std::vector<Molecule> keep = getMolecules();
    for (int d = 0; d < 3; ++d) {
        _bbMin[d] = bBoxMin[d];
        _bbMax[d] = bBoxMax[d];
    }
    makeContext(1);
    _pending = keep;
    _uploaded = false;
    int32_t cpd[3];
    double cl[3];
    check(mdx_cell_geometry(_ctx, cpd, cl), "mdx_cell_geometry");
    // sendParticlesTogether: false when a dimension holds fewer than two
    // inner cells (LinkedCells.cpp:186-193)
    return cpd[0] >= 4 && cpd[1] >= 4 && cpd[2] >= 4;
}

void HIPLinkedCells::addParticles(std::vector<Molecule> &particles, bool) {
    if (_uploaded) {
        std::vector<Molecule> cur = getMolecules();
        _pending.insert(_pending.begin(), cur.begin(), cur.end());
        _uploaded = false;
    }
    _pending.insert(_pending.end(), particles.begin(), particles.end());
}

void HIPLinkedCells::setThermostat(double targetT, bool lrc, bool explosionHeuristics) {
    _haveThermostat = true;
    _targetT = targetT;
    _lrc = lrc;
    _heuristics = explosionHeuristics;
    check(mdx_set_thermostat(_ctx, 1, &_targetT, _lrc ? 1 : 0, _heuristics ? 1 : 0),
          "mdx_set_thermostat");
}

void HIPLinkedCells::setBoundaries(const int32_t types[3]) {
    _haveBoundaries = true;
    for (int d = 0; d < 3; ++d) _boundaries[d] = types[d];
    check(mdx_set_boundaries(_ctx, _boundaries), "mdx_set_boundaries");
}

void HIPLinkedCells::update() {
    if (!_uploaded) {
        const size_t n = _pending.size();
        std::vector<double> r(3 * n), v(3 * n), q(4 * n), L(3 * n);
        std::vector<int32_t> cid(n);
        std::vector<uint64_t> ids(n);
        for (size_t i = 0; i < n; ++i) {
            const Molecule &m = _pending[i];
            for (int d = 0; d < 3; ++d) {
                r[3 * i + d] = m.r[d];
                v[3 * i + d] = m.v[d];
                L[3 * i + d] = m.D[d];
            }
            for (int k = 0; k < 4; ++k) q[4 * i + k] = m.q[k];
            cid[i] = m.cid;
            ids[i] = m.id;
        }
        check(mdx_upload(_ctx, (int64_t)n, r.data(), v.data(), q.data(), L.data(), cid.data(),
                         ids.data()),
              "mdx_upload");
        _n = n;
        _cids = cid;
        _pending.clear();
        _uploaded = true;
    }
    check(mdx_update_container(_ctx), "mdx_update_container");
}

void HIPLinkedCells::traverseCells(CellProcessor &cellProcessor) {
    auto *hip = dynamic_cast<HIPCellProcessor *>(&cellProcessor);
    if (!hip)
        die("HIPLinkedCells traverses on the device; only HIPCellProcessor can be applied "
            "(other cell processors need ls1's host LinkedCells)");
    if (!_uploaded) update();
    cellProcessor.initTraversal();
    hip->traverseOnDevice(*this);
    cellProcessor.endTraversal();
}

unsigned long HIPLinkedCells::getNumberOfParticles() {
    return _uploaded ? _n : (unsigned long)_pending.size();
}

void HIPLinkedCells::clear() {
    _pending.clear();
    _uploaded = false;
    _n = 0;
}

std::vector<Molecule> HIPLinkedCells::getMolecules() {
    if (!_uploaded) return _pending;
    const size_t n = _n;
    std::vector<double> r(3 * n), v(3 * n), q(4 * n), L(3 * n), F(3 * n), M(3 * n), Vi(3 * n);
    std::vector<uint64_t> ids(n);
    check(mdx_download(_ctx, (int64_t)n, r.data(), v.data(), q.data(), L.data(), F.data(), M.data(),
                       Vi.data(), ids.data()),
          "mdx_download");
    std::vector<Molecule> out(n);
    // download preserves the upload order, so component ids come from it
    for (size_t i = 0; i < n; ++i) {
        Molecule &m = out[i];
        m.id = ids[i];
        m.cid = _cids[i];
        for (int d = 0; d < 3; ++d) {
            m.r[d] = r[3 * i + d];
            m.v[d] = v[3 * i + d];
            m.D[d] = L[3 * i + d];
            m.F[d] = F[3 * i + d];
            m.M[d] = M[3 * i + d];
            m.Vi[d] = Vi[3 * i + d];
        }
        for (int k = 0; k < 4; ++k) m.q[k] = q[4 * i + k];
    }
    return out;
}

void HIPLinkedCells::setPrecision(int precision) {
    check(mdx_set_precision(_ctx, precision), "mdx_set_precision");
    _precision = precision;
}

std::vector<int64_t> HIPLinkedCells::getCellIndices() {
    std::vector<int64_t> out(_n);
    check(mdx_cell_ids(_ctx, (int64_t)_n, out.data()), "mdx_cell_ids");
    return out;
}

// ------------------------------------------------------------- processor

HIPCellProcessor::HIPCellProcessor(Domain &domain, double cutoffRadius, double LJcutoffRadius)
    : CellProcessor(cutoffRadius, LJcutoffRadius), _domain(domain) {}

// VectorizedCellProcessor::initTraversal (VectorizedCellProcessor.cpp:109-120)
void HIPCellProcessor::initTraversal() {
    for (double &m : _macro) m = 0.0;
    _haveDevice = false;
    _hostVisits = 0;
}

void HIPCellProcessor::processCellPair(ParticleCell &c1, ParticleCell &c2, bool) {
    if (c1.getMoleculeCount() == 0 || c2.getMoleculeCount() == 0) return;
    if (c1.isHaloCell() && c2.isHaloCell()) return;
    ++_hostVisits;
}

void HIPCellProcessor::processCell(ParticleCell &c) {
    if (c.isHaloCell() || c.getMoleculeCount() < 2) return;
    ++_hostVisits;
}

double HIPCellProcessor::processSingleMolecule(Molecule *, ParticleCell &) {
    throw std::runtime_error("HIPCellProcessor::processSingleMolecule is not supported "
                             "(the reference's VectorizedCellProcessor does not support it either)");
}

void HIPCellProcessor::traverseOnDevice(HIPLinkedCells &cells) {
    if (mdx_traverse(cells.context(), _macro) != MDX_OK)
        throw std::runtime_error(std::string("mdx_traverse failed: ") +
                                 mdx_last_error(cells.context()));
    _haveDevice = true;
}

// VectorizedCellProcessor::endTraversal (VectorizedCellProcessor.cpp:122-155)
void HIPCellProcessor::endTraversal() {
    _domain.setLocalVirial(_macro[MDX_VIRIAL] + 3.0 * _macro[MDX_MYRF]);
    _domain.setLocalUpot(_macro[MDX_UPOT6LJ] / 6.0 + _macro[MDX_UPOTXPOLES] + _macro[MDX_MYRF]);
}

// -------------------------------------------------------------- Leapfrog

void HIPLeapfrog::eventNewTimestep(ParticleContainer *molCont, Domain *) {
    auto *c = dynamic_cast<HIPLinkedCells *>(molCont);
    if (!c) throw std::runtime_error("HIPLeapfrog needs a HIPLinkedCells container");
    if (mdx_leapfrog_pre(c->context(), _timestepLength) != MDX_OK)
        throw std::runtime_error(std::string("mdx_leapfrog_pre: ") + mdx_last_error(c->context()));
}

void HIPLeapfrog::eventForcesCalculated(ParticleContainer *molCont, Domain *domain) {
    auto *c = dynamic_cast<HIPLinkedCells *>(molCont);
    if (!c) throw std::runtime_error("HIPLeapfrog needs a HIPLinkedCells container");
    double sums[4];
    if (mdx_leapfrog_post(c->context(), _timestepLength, sums) != MDX_OK)
        throw std::runtime_error(std::string("mdx_leapfrog_post: ") + mdx_last_error(c->context()));
    domain->setLocalSums(sums[0], sums[1], (uint64_t)sums[2], (uint64_t)sums[3]);
    // Domain::calculateGlobalValues and the velocity scaling ran on the device
    // (when mdx_set_thermostat enabled them); mirror their results
    double th[8];
    if (mdx_thermostat_state(c->context(), th) != MDX_OK)
        throw std::runtime_error(std::string("mdx_thermostat_state: ") + mdx_last_error(c->context()));
    domain->setUpotCorr(th[1]);
    domain->setVirialCorr(th[2]);
    domain->setGlobalBetas(th[5], th[6]);
}

}  // namespace ls1amd
