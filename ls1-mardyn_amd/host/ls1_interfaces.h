// ls1_interfaces.h -- the host-side interfaces of ls1-MarDyn that the HIP
// path plugs into, restated with the same names, virtuals and argument
// meaning so the device implementation is a drop-in:
//   CellProcessor      src/particleContainer/adapter/CellProcessor.h:29-94
//   ParticleContainer  src/particleContainer/ParticleContainer.h:85-279 (the
//                      members the force path uses)
//   Domain             src/Domain.h (local Upot / virial sinks, global values
//                      of Domain::calculateGlobalValues, Domain.cpp:157-270)
//   Integrator         src/integrators/Integrator.h:54,66
// Errors follow the reference convention: MARDYN_EXIT-style aborts
// (src/utils/mardyn_assert.h:21-40) become std::runtime_error here.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ls1amd {

// One molecule as the reference's ICRVQD phase-space format carries it
// (src/io/ASCIIReader.cpp:370-400) plus the force outputs of calcFM.
struct Molecule {
    uint64_t id = 0;
    int cid = 0;  // 0-based component id
    double r[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    double q[4] = {1, 0, 0, 0};
    double D[3] = {0, 0, 0};  // angular momentum (lab frame)
    double F[3] = {0, 0, 0}, M[3] = {0, 0, 0}, Vi[3] = {0, 0, 0};
};

// The cell handle CellProcessor receives (FullParticleCell's role). The
// device container exposes its cells through this interface; molecule data
// stay on the device.
class ParticleCell {
public:
    virtual ~ParticleCell() = default;
    virtual bool isHaloCell() const = 0;
    virtual long getCellIndex() const = 0;
    virtual int getMoleculeCount() const = 0;
};

class CellProcessor {
public:
    CellProcessor(double cutoffRadius, double LJcutoffRadius)
        : _cutoffRadius(cutoffRadius), _LJCutoffRadius(LJcutoffRadius) {}
    virtual ~CellProcessor() = default;
    virtual void initTraversal() = 0;
    virtual void preprocessCell(ParticleCell &cell) = 0;
    virtual void processCellPair(ParticleCell &cell1, ParticleCell &cell2, bool sumAll = false) = 0;
    virtual void processCell(ParticleCell &cell) = 0;
    virtual double processSingleMolecule(Molecule *m1, ParticleCell &cell2) = 0;
    virtual void postprocessCell(ParticleCell &cell) = 0;
    virtual void endTraversal() = 0;
    double getCutoffRadius() const { return _cutoffRadius; }
    double getLJCutoffRadius() const { return _LJCutoffRadius; }

protected:
    double _cutoffRadius, _LJCutoffRadius;
};

// Macroscopic sinks and global values (Domain.cpp:115-119, 142-146, 157-270;
// single process, one global thermostat). The long-range corrections and the
// thermostat's factors are computed on the device inside the step and set
// here by the integrator adapter.
class Domain {
public:
    explicit Domain(const double box[3]) {
        for (int d = 0; d < 3; ++d) _box[d] = box[d];
    }
    void setLocalUpot(double u) { _localUpot = u; }
    void setLocalVirial(double w) { _localVirial = w; }
    double getLocalUpot() const { return _localUpot; }
    double getLocalVirial() const { return _localVirial; }
    double getGlobalLength(int d) const { return _box[d]; }
    void setglobalNumMolecules(uint64_t n) { _n = n; }
    void setLocalSums(double summv2, double sumIw2, uint64_t n, uint64_t rotDOF) {
        _summv2 = summv2;
        _sumIw2 = sumIw2;
        _n = n;
        _rotDOF = rotDOF;
    }
    // T = (sum m v^2 + sum I w^2) / (3 N + rotDOF)   (Domain.cpp:240-242)
    double getGlobalTemperature() const {
        double iw2 = _rotDOF > 0 ? _sumIw2 : 0.0;
        return _n ? (_summv2 + iw2) / (double)(3 * _n + _rotDOF) : 0.0;
    }
    double getGlobalRho() const { return _n / (_box[0] * _box[1] * _box[2]); }
    // p = rho T + rho <virial> / 3   (Domain::getGlobalPressure, Domain.cpp:142-146)
    double getGlobalPressure() const {
        double rho = getGlobalRho();
        return getGlobalTemperature() * rho + rho * (_n ? getGlobalVirial() / _n : 0.0) / 3.;
    }
    // Domain::setUpotCorr / setVirialCorr; global = local + correction
    // (Domain.cpp:187-188)
    void setUpotCorr(double u) { _upotCorr = u; }
    void setVirialCorr(double w) { _virialCorr = w; }
    double getGlobalUpot() const { return _localUpot + _upotCorr; }
    double getGlobalVirial() const { return _localVirial + _virialCorr; }
    double getAverageGlobalUpot() const { return _n ? getGlobalUpot() / _n : 0.0; }
    // the global thermostat's factors of the last step (Domain.cpp:247-259)
    void setGlobalBetas(double bt, double br) { _betaTrans = bt; _betaRot = br; }
    double getGlobalBetaTrans() const { return _betaTrans; }
    double getGlobalBetaRot() const { return _betaRot; }

private:
    double _box[3];
    double _localUpot = 0, _localVirial = 0, _summv2 = 0, _sumIw2 = 0;
    double _upotCorr = 0, _virialCorr = 0, _betaTrans = 1, _betaRot = 1;
    uint64_t _n = 0, _rotDOF = 0;
};

class ParticleContainer {
public:
    virtual ~ParticleContainer() = default;
    virtual bool rebuild(double bBoxMin[3], double bBoxMax[3]) = 0;
    virtual void update() = 0;
    virtual void addParticles(std::vector<Molecule> &particles, bool checkWhetherDuplicate = false) = 0;
    virtual void traverseCells(CellProcessor &cellProcessor) = 0;
    virtual unsigned long getNumberOfParticles() = 0;
    virtual void clear() = 0;
    virtual void deleteOuterParticles() = 0;
    virtual double getCutoff() const = 0;
    virtual double getBoundingBoxMin(int d) const = 0;
    virtual double getBoundingBoxMax(int d) const = 0;
    virtual void updateMoleculeCaches() = 0;
    // copies of the molecules (the device keeps the state)
    virtual std::vector<Molecule> getMolecules() = 0;
};

class Integrator {
public:
    explicit Integrator(double dt) : _timestepLength(dt) {}
    virtual ~Integrator() = default;
    virtual void eventNewTimestep(ParticleContainer *molCont, Domain *domain) = 0;
    virtual void eventForcesCalculated(ParticleContainer *molCont, Domain *domain) = 0;
    double getTimestepLength() const { return _timestepLength; }

protected:
    double _timestepLength;
};

}  // namespace ls1amd
