// HIPLinkedCellsAdapter.h -- the binding a maintainer adds to ls1-MarDyn's
// source tree to run the short-range force path on one MI355X through the C
// ABI of include/ls1amd.h. Written against ls1's own headers (compile with
// -I<ls1>/src -I<ls1>/libs/rapidxml and ls1's precision macro, e.g.
// -DMARDYN_DPDP); nothing here is used by the library itself.
//
//   HIPLinkedCellsAdapter   : ParticleContainer (src/particleContainer/
//                             ParticleContainer.h:85-279) whose molecules,
//                             cells and SoA live on the device. The host
//                             keeps a mirror of ParticleCells on ls1's halo-
//                             extended cell grid for iterators, plugins and
//                             I/O: it is refreshed from the device when the
//                             host asks for molecules after device work, and
//                             uploaded again before the next device step when
//                             the host may have changed it.
//   HIPCellProcessorAdapter : CellProcessor (src/particleContainer/adapter/
//                             CellProcessor.h:29-94) with
//                             VectorizedCellProcessor's physics and
//                             endTraversal accounting
//                             (VectorizedCellProcessor.cpp:122-155).
//   HIPLeapfrogAdapter      : Integrator (src/integrators/Integrator.h:34-80)
//                             with Leapfrog's two transitions
//                             (Leapfrog.cpp:32-156) on the device.
//
// Selection: <datastructure type="HIPLinkedCells"> in Simulation::readXML
// (Simulation.cpp:471-519), the processor in Simulation::prepare_start
// (Simulation.cpp:851-862), <integrator type="HIPLeapfrog">.
// Simulation::updateForces (calcFM, Simulation.cpp:835-845) is skipped for
// this container: the device computes molecule forces, torques and virials
// inside the traversal, and the mirror carries them.
//
// XML (inside <datastructure type="HIPLinkedCells">):
//   <cellsInCutoffRadius>1</cellsInCutoffRadius>   (only 1, as LinkedCells)
//   <device>0</device>
//   <precision>FP64 | FP32</precision>             (FP32 = FP32 pair
//                                                   arithmetic, FP64 sums)
// Errors follow ls1's convention: every failing mdx_* call ends in
// MARDYN_EXIT with mdx_last_error's message (src/utils/mardyn_assert.h:21-40).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <variant>
#include <vector>

#include "integrators/Integrator.h"
#include "ls1amd.h"
#include "particleContainer/ParticleContainer.h"
#include "particleContainer/adapter/CellProcessor.h"

class Domain;

class HIPLinkedCellsAdapter : public ParticleContainer {
public:
	HIPLinkedCellsAdapter();
	HIPLinkedCellsAdapter(double bBoxMin[3], double bBoxMax[3], double cutoffRadius);
	~HIPLinkedCellsAdapter() override;

	void readXML(XMLfileUnits& xmlconfig) override;
	bool rebuild(double bBoxMin[3], double bBoxMax[3]) override;
	void update() override;
	bool addParticle(Molecule& particle, bool inBoxCheckedAlready = false, bool checkWhetherDuplicate = false,
					 const bool& rebuildCaches = false) override;
	void addParticles(std::vector<Molecule>& particles, bool checkWhetherDuplicate = false) override;
	void traverseCells(CellProcessor& cellProcessor) override;
	void traverseNonInnermostCells(CellProcessor& cellProcessor) override;
	void traversePartialInnermostCells(CellProcessor& cellProcessor, unsigned int stage, int stageCount) override;
	ParticleIterator iterator(ParticleIterator::Type t) override;
	RegionParticleIterator regionIterator(const double startCorner[3], const double endCorner[3],
										  ParticleIterator::Type t) override;
	unsigned long getNumberOfParticles(ParticleIterator::Type t = ParticleIterator::ALL_CELLS) override;
	void clear() override;
	void deleteOuterParticles() override;
	double getHaloWidthForDimension(int index) const override;
	double getCutoff() const override { return _cutoffRadius; }
	void setCutoff(double rc) override { _cutoffRadius = rc; }
	void deleteMolecule(ParticleIterator& moleculeIter, const bool& rebuildCaches) override;
	double getEnergy(ParticlePairsHandler* particlePairsHandler, Molecule* m1, CellProcessor& cellProcessor) override;
	void updateInnerMoleculeCaches() override {}
	void updateBoundaryAndHaloMoleculeCaches() override {}
	void updateMoleculeCaches() override {}
	std::variant<ParticleIterator, SingleCellIterator<ParticleCell>> getMoleculeAtPosition(const double pos[3]) override;
	unsigned long initCubicGrid(std::array<unsigned long, 3> numMoleculesPerDimension,
								std::array<double, 3> simBoxLength) override;
	double* getCellLength() override { return _cellLength; }
	std::string getConfigurationAsString() override;
	std::vector<unsigned long> getParticleCellStatistics() override;

	// MemoryProfilable
	size_t getTotalSize() override;
	void printSubInfo(int offset) override;
	std::string getName() override { return "HIPLinkedCells"; }

	// The device side, for the adapters below.
	mdx_context* context() { return _ctx; }
	void setLJCutoff(double rcLJ) { _LJCutoffRadius = rcLJ; }
	// Uploads the mirror if the host may have changed it since the last
	// upload; every device operation calls it first.
	void syncToDevice();
	// The device has changed the state: the mirror is stale.
	void deviceChanged() { _hostValid = false; }

private:
	void check(int rc, const char* what) const;
	void configureDevice();
	void syncToHost();
	void initializeMirror();
	long cellIndexOfPoint(const double p[3]) const;

	mdx_context* _ctx = nullptr;
	int _device = 0;
	int _precision = MDX_PREC_FP64;
	int _cellsInCutoff = 1;
	double _cutoffRadius = 0.0, _LJCutoffRadius = 0.0;
	int _cellsPerDimension[3] = {0, 0, 0};
	double _cellLength[3] = {0, 0, 0};
	double _haloBoundingBoxMin[3] = {0, 0, 0}, _haloBoundingBoxMax[3] = {0, 0, 0};
	int _haloWidthInNumCells[3] = {1, 1, 1};
	// host mirror on ls1's halo-extended cell grid (halo cells stay empty:
	// the device uses image shifts instead of halo copies)
	std::vector<ParticleCell> _cells;
	bool _hostValid = true;   // the mirror holds the device state
	bool _hostDirty = false;  // the host may have changed the mirror
	bool _domainSet = false;
	unsigned long _numUploaded = 0;
};

class HIPCellProcessorAdapter : public CellProcessor {
public:
	HIPCellProcessorAdapter(Domain& domain, double cutoffRadius, double LJcutoffRadius);

	void initTraversal() override;
	// The device owns the pair enumeration: a host cell walk driving this
	// processor is refused (HIPLinkedCellsAdapter never issues one).
	void preprocessCell(ParticleCell&) override {}
	void processCellPair(ParticleCell& cell1, ParticleCell& cell2, bool sumAll = false) override;
	void processCell(ParticleCell& cell) override;
	double processSingleMolecule(Molecule* m1, ParticleCell& cell2) override;
	void postprocessCell(ParticleCell&) override {}
	void endTraversal() override;

	// called by HIPLinkedCellsAdapter::traverseCells between initTraversal
	// and endTraversal
	void traverseOnDevice(HIPLinkedCellsAdapter& container);

private:
	Domain& _domain;
	double _macro[4] = {0, 0, 0, 0};
};

class HIPLeapfrogAdapter : public Integrator {
public:
	HIPLeapfrogAdapter() = default;
	explicit HIPLeapfrogAdapter(double timestepLength) : Integrator(timestepLength) {}

	void readXML(XMLfileUnits& xmlconfig) override;
	void init() override {}
	// Leapfrog::transition2to3 (Leapfrog.cpp:63-147): upd_postF on the
	// device, the sums to the global thermostat's Domain entries
	void eventForcesCalculated(ParticleContainer* moleculeContainer, Domain* domain) override;
	// Leapfrog::transition1to2 (Leapfrog.cpp:45-61): upd_preF on the device
	void eventNewTimestep(ParticleContainer* moleculeContainer, Domain* domain) override;

private:
	static HIPLinkedCellsAdapter& device(ParticleContainer* moleculeContainer);
};
