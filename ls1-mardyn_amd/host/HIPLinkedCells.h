// HIPLinkedCells.h -- the device-resident drop-ins behind ls1's interfaces:
//   HIPLinkedCells   : ParticleContainer whose cells, molecules and SoA live
//                      in HBM (replaces LinkedCells, LinkedCells.cpp); its
//                      traverseCells runs the whole C08-equivalent traversal
//                      on the device.
//   HIPCellProcessor : CellProcessor with VectorizedCellProcessor's physics
//                      and macroscopic accounting (VectorizedCellProcessor.cpp);
//                      endTraversal publishes Upot / virial to the Domain
//                      exactly like VectorizedCellProcessor::endTraversal.
//   HIPLeapfrog      : Integrator with Leapfrog's two transitions
//                      (Leapfrog.cpp:32-156) on the device.
// All three talk to the GPU only through the C ABI of include/ls1amd.h.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/ls1amd.h"
#include "ls1_interfaces.h"

namespace ls1amd {

// Component records and mixing as ls1's ASCII header carries them (see
// include/ls1amd.h for the layout).
struct ComponentModel {
    int ncomp = 0;
    std::vector<double> desc;  // flattened records
    std::vector<double> mix;   // (xi, eta) per component pair i < j
    double epsilonRF = 1.0;
};

class HIPCellProcessor;

class HIPLinkedCells : public ParticleContainer {
public:
    // LinkedCells(bBoxMin, bBoxMax, cutoffRadius) (LinkedCells.cpp:38-106);
    // the component model replaces the global Simulation lookup the reference
    // makes. periodic selects the boundary handling of
    // DomainDecompBase::exchangeMolecules (true) or the open boundaries of
    // the reference's test set-up (false).
    HIPLinkedCells(double bBoxMin[3], double bBoxMax[3], double cutoffRadius, double LJcutoffRadius,
                   const ComponentModel &model, bool periodic = true, int device = 0,
                   int precision = MDX_PREC_FP64);
    ~HIPLinkedCells() override;

    bool rebuild(double bBoxMin[3], double bBoxMax[3]) override;
    void update() override;
    void addParticles(std::vector<Molecule> &particles, bool checkWhetherDuplicate = false) override;
    void traverseCells(CellProcessor &cellProcessor) override;
    unsigned long getNumberOfParticles() override;
    void clear() override;
    void deleteOuterParticles() override {}  // no halo copies exist
    double getCutoff() const override { return _cutoff; }
    double getBoundingBoxMin(int d) const override { return _bbMin[d]; }
    double getBoundingBoxMax(int d) const override { return _bbMax[d]; }
    void updateMoleculeCaches() override {}  // built by update()
    std::vector<Molecule> getMolecules() override;

    mdx_context *context() { return _ctx; }
    void setPrecision(int precision);
    int getDevice() const { return _device; }
    int getPrecision() const { return _precision; }
    // the global velocity-scaling thermostat and the homogeneous long-range
    // correction (mdx_set_thermostat), kept over rebuild()
    void setThermostat(double targetT, bool lrc, bool explosionHeuristics = true);
    // boundary type per dimension (mdx_set_boundaries), kept over rebuild()
    void setBoundaries(const int32_t types[3]);
    // ls1 cell index of every molecule, in insertion order
    std::vector<int64_t> getCellIndices();

private:
    void check(int rc, const char *what);
    void setDomain(int cellMode);
    void makeContext(int cellMode);
    mdx_context *_ctx = nullptr;
    ComponentModel _model;
    double _bbMin[3], _bbMax[3], _cutoff, _ljCutoff;
    bool _periodic;
    int _device, _precision;
    bool _haveThermostat = false, _lrc = false, _heuristics = true, _haveBoundaries = false;
    double _targetT = 0.0;
    int32_t _boundaries[3] = {MDX_BOUNDARY_PERIODIC, MDX_BOUNDARY_PERIODIC, MDX_BOUNDARY_PERIODIC};
    std::vector<Molecule> _pending;
    std::vector<int32_t> _cids;  // component ids in upload order
    bool _uploaded = false;
    unsigned long _n = 0;
    friend class HIPCellProcessor;
};

class HIPCellProcessor : public CellProcessor {
public:
    HIPCellProcessor(Domain &domain, double cutoffRadius, double LJcutoffRadius);
    void initTraversal() override;
    // The device owns the pair enumeration: when the traversal is driven by
    // a host-side cell walk these calls only count visits (the role of
    // LinkedCellsTest's CellProcessorStub); HIPLinkedCells::traverseCells
    // never issues them.
    void preprocessCell(ParticleCell &) override {}
    void processCellPair(ParticleCell &c1, ParticleCell &c2, bool sumAll = false) override;
    void processCell(ParticleCell &c) override;
    double processSingleMolecule(Molecule *, ParticleCell &) override;
    void postprocessCell(ParticleCell &) override {}
    void endTraversal() override;

    // called by HIPLinkedCells::traverseCells between init/endTraversal
    void traverseOnDevice(HIPLinkedCells &cells);
    const double *macro() const { return _macro; }
    long long hostVisits() const { return _hostVisits; }

private:
    Domain &_domain;
    double _macro[4] = {0, 0, 0, 0};
    bool _haveDevice = false;
    long long _hostVisits = 0;
};

class HIPLeapfrog : public Integrator {
public:
    explicit HIPLeapfrog(double dt) : Integrator(dt) {}
    void eventNewTimestep(ParticleContainer *molCont, Domain *domain) override;
    void eventForcesCalculated(ParticleContainer *molCont, Domain *domain) override;
};

}  // namespace ls1amd
