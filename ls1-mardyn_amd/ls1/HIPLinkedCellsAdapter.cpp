// HIPLinkedCellsAdapter.cpp -- see HIPLinkedCellsAdapter.h.
#include "HIPLinkedCellsAdapter.h"

#include <algorithm>
#include <cmath>
#include <sstream>

#include "Domain.h"
#include "Simulation.h"
#include "ensemble/EnsembleBase.h"
#include "molecules/Comp2Param.h"
#include "molecules/Component.h"
#include "molecules/Molecule.h"
#include "molecules/mixingrules/LorentzBerthelot.h"
#include "utils/Logger.h"
#include "utils/mardyn_assert.h"
#include "utils/xmlfileUnits.h"
#include "WrapOpenMP.h"

namespace {

// Component records in the layout include/ls1amd.h documents (the ASCII
// phase-space header, ASCIIReader.cpp:152-232): counts, LJ sites, charges,
// dipoles, quadrupoles, principal moments.
std::vector<double> flattenComponents(const std::vector<Component>& comps, double rcLJ) {
	std::vector<double> d;
	for (const Component& c : comps) {
		d.insert(d.end(), {double(c.numLJcenters()), double(c.numCharges()), double(c.numDipoles()),
						   double(c.numQuadrupoles())});
		for (unsigned i = 0; i < c.numLJcenters(); ++i) {
			const LJcenter& s = c.ljcenter(i);
			d.insert(d.end(), {s.rx(), s.ry(), s.rz(), s.m(), s.eps(), s.sigma(), rcLJ,
							   s.shiftRequested() ? 1.0 : 0.0});
		}
		for (unsigned i = 0; i < c.numCharges(); ++i) {
			const Charge& s = c.charge(i);
			d.insert(d.end(), {s.rx(), s.ry(), s.rz(), s.m(), s.q()});
		}
		for (unsigned i = 0; i < c.numDipoles(); ++i) {
			const Dipole& s = c.dipole(i);
			d.insert(d.end(), {s.rx(), s.ry(), s.rz(), s.ex(), s.ey(), s.ez(), s.absMy()});
		}
		for (unsigned i = 0; i < c.numQuadrupoles(); ++i) {
			const Quadrupole& s = c.quadrupole(i);
			d.insert(d.end(), {s.rx(), s.ry(), s.rz(), s.ex(), s.ey(), s.ez(), s.absQ()});
		}
		d.insert(d.end(), {c.I11(), c.I22(), c.I33()});
	}
	return d;
}

// (xi, eta) per component pair i < j from the ensemble's mixing rules
// (EnsembleBase.h:124-136); pairs without a rule mix with (1, 1).
std::vector<double> mixingCoefficients(Ensemble& ensemble, size_t ncomp) {
	std::vector<double> mix;
	auto& rules = ensemble.getMixingrules();
	for (size_t i = 0; i < ncomp; ++i)
		for (size_t j = i + 1; j < ncomp; ++j) {
			double xi = 1.0, eta = 1.0;
			auto a = rules.find((int)i);
			if (a != rules.end()) {
				auto b = a->second.find((int)j);
				if (b != a->second.end() && b->second) {
					auto* lb = dynamic_cast<LorentzBerthelotMixingRule*>(b->second.get());
					if (lb) {
						xi = lb->getXi();
						eta = lb->getEta();
					}
				}
			}
			mix.push_back(xi);
			mix.push_back(eta);
		}
	return mix;
}

// The exact tables VectorizedCellProcessor builds from Comp2Param
// (VectorizedCellProcessor.cpp:40-80): 24 eps, sigma^2 and 6 * shift over
// all LJ centres, read from the same parameter streams.
void comp2ParamTables(Domain& domain, const std::vector<Component>& comps, std::vector<double>& eps24,
					  std::vector<double>& sig2, std::vector<double>& shift6) {
	std::vector<size_t> base(comps.size(), 0);
	size_t n = 0;
	for (size_t c = 0; c < comps.size(); ++c) {
		base[c] = n;
		n += comps[c].numLJcenters();
	}
	eps24.assign(n * n, 0.0);
	sig2.assign(n * n, 0.0);
	shift6.assign(n * n, 0.0);
	for (size_t ci = 0; ci < comps.size(); ++ci)
		for (size_t cj = 0; cj < comps.size(); ++cj) {
			ParaStrm& p = domain.getComp2Params()(comps[ci].ID(), comps[cj].ID());
			p.reset_read();
			for (size_t a = 0; a < comps[ci].numLJcenters(); ++a)
				for (size_t b = 0; b < comps[cj].numLJcenters(); ++b) {
					double e, s, sh;
					p >> e;
					p >> s;
					p >> sh;
					const size_t k = (base[ci] + a) * n + base[cj] + b;
					eps24[k] = e;
					sig2[k] = s;
					shift6[k] = sh;
				}
		}
}

}  // namespace

// ------------------------------------------------------- HIPLinkedCellsAdapter

HIPLinkedCellsAdapter::HIPLinkedCellsAdapter() = default;

HIPLinkedCellsAdapter::HIPLinkedCellsAdapter(double bBoxMin[3], double bBoxMax[3], double cutoffRadius)
	: ParticleContainer(bBoxMin, bBoxMax), _cutoffRadius(cutoffRadius), _LJCutoffRadius(cutoffRadius) {
	rebuild(bBoxMin, bBoxMax);
}

HIPLinkedCellsAdapter::~HIPLinkedCellsAdapter() {
	if (_ctx) mdx_destroy(_ctx);
}

void HIPLinkedCellsAdapter::check(int rc, const char* what) const {
	if (rc != MDX_OK) {
		std::ostringstream msg;
		msg << "HIPLinkedCells: " << what << " failed (" << rc << "): " << mdx_last_error(_ctx) << std::endl;
		MARDYN_EXIT(msg.str());
	}
}

void HIPLinkedCellsAdapter::readXML(XMLfileUnits& xmlconfig) {
	// LinkedCells::readXML (LinkedCells.cpp:124-130); the traversal choice has
	// no meaning here (the device owns the enumeration)
	_cellsInCutoff = xmlconfig.getNodeValue_int("cellsInCutoffRadius", 1);
	if (_cellsInCutoff != 1) MARDYN_EXIT("HIPLinkedCells: cellsInCutoffRadius must be 1");
	_device = xmlconfig.getNodeValue_int("device", 0);
	std::string prec = "FP64";
	xmlconfig.getNodeValue("precision", prec);
	if (prec == "FP64")
		_precision = MDX_PREC_FP64;
	else if (prec == "FP32")
		_precision = MDX_PREC_FP32_FP64ACC;
	else
		MARDYN_EXIT("HIPLinkedCells: precision must be FP64 or FP32, not " + prec);
	Log::global_log->info() << "HIPLinkedCells: device " << _device << ", precision " << prec << std::endl;
}

// Component model, exact VCP tables, domain and precision on a fresh context
// (LinkedCells::rebuild, LinkedCells.cpp:132-201: float cell edge).
void HIPLinkedCellsAdapter::configureDevice() {
	if (_ctx) check(mdx_destroy(_ctx), "mdx_destroy");
	_ctx = nullptr;
	check(mdx_create(_device, &_ctx), "mdx_create");
	if (global_simulation) {
		_cutoffRadius = global_simulation->getcutoffRadius();
		_LJCutoffRadius = global_simulation->getLJCutoff();
	}
	auto& comps = *global_simulation->getEnsemble()->getComponents();
	std::vector<double> desc = flattenComponents(comps, _LJCutoffRadius);
	std::vector<double> mix = mixingCoefficients(*global_simulation->getEnsemble(), comps.size());
	Domain& domain = *global_simulation->getDomain();
	check(mdx_set_components(_ctx, (int)comps.size(), desc.data(), (int)desc.size(),
							 mix.empty() ? nullptr : mix.data(), domain.getepsilonRF()),
		  "mdx_set_components");
	std::vector<double> eps24, sig2, shift6;
	comp2ParamTables(domain, comps, eps24, sig2, shift6);
	if (!eps24.empty())
		check(mdx_set_lj_tables(_ctx, (int)std::lround(std::sqrt((double)eps24.size())), eps24.data(),
								sig2.data(), shift6.data()),
			  "mdx_set_lj_tables");
	double box[3];
	for (int d = 0; d < 3; ++d) {
		if (_boundingBoxMin[d] != 0.0)
			MARDYN_EXIT("HIPLinkedCells: the domain must start at the origin (one process)");
		box[d] = _boundingBoxMax[d] - _boundingBoxMin[d];
	}
	check(mdx_set_domain(_ctx, box, _cutoffRadius, _LJCutoffRadius, /*periodic=*/1, /*cell_mode=*/1),
		  "mdx_set_domain");
	check(mdx_set_precision(_ctx, _precision), "mdx_set_precision");
	int32_t cpd[3];
	check(mdx_cell_geometry(_ctx, cpd, _cellLength), "mdx_cell_geometry");
	for (int d = 0; d < 3; ++d) {
		_cellsPerDimension[d] = cpd[d];
		_haloBoundingBoxMin[d] = _boundingBoxMin[d] - _cellLength[d];
		_haloBoundingBoxMax[d] = _boundingBoxMax[d] + _cellLength[d];
	}
	_domainSet = true;
}

bool HIPLinkedCellsAdapter::rebuild(double bBoxMin[3], double bBoxMax[3]) {
	// keep the molecules over a rebuild (the mirror is re-binned below)
	syncToHost();
	std::vector<Molecule> keep;
	for (auto& cell : _cells)
		for (auto it = cell.iterator(); it.isValid(); ++it) keep.push_back(*it);
	for (int d = 0; d < 3; ++d) {
		_boundingBoxMin[d] = bBoxMin[d];
		_boundingBoxMax[d] = bBoxMax[d];
	}
	configureDevice();
	initializeMirror();
	for (auto& m : keep) addParticle(m, false, false, false);
	_hostDirty = true;
	// LinkedCells::rebuild returns sendParticlesTogether (LinkedCells.cpp:
	// 186-193): false when a dimension holds fewer than 2 inner cells
	return _cellsPerDimension[0] >= 4 && _cellsPerDimension[1] >= 4 && _cellsPerDimension[2] >= 4;
}

// The mirror grid and the cell flags, as LinkedCells::initializeCells
// (LinkedCells.cpp:689-720) sets them up.
void HIPLinkedCellsAdapter::initializeMirror() {
	long ncell = (long)_cellsPerDimension[0] * _cellsPerDimension[1] * _cellsPerDimension[2];
	_cells.clear();
	_cells.resize(ncell);
	ParticleCell::_cellBorderAndFlagManager.init(_cellsPerDimension, _haloBoundingBoxMin, _haloBoundingBoxMax,
												 _boundingBoxMin, _boundingBoxMax, _cellLength,
												 _haloWidthInNumCells);
	for (int iz = 0; iz < _cellsPerDimension[2]; ++iz)
		for (int iy = 0; iy < _cellsPerDimension[1]; ++iy)
			for (int ix = 0; ix < _cellsPerDimension[0]; ++ix) {
				const long idx = ((long)iz * _cellsPerDimension[1] + iy) * _cellsPerDimension[0] + ix;
				// the cell's box and halo flag follow from its index through
				// the flag manager initialised above
				_cells[idx].setCellIndex(idx);
			}
	_hostValid = true;
}

// LinkedCells::getCellIndexOfPoint's first estimate (LinkedCells.cpp:870-
// 905), clamped to the grid; the mirror only needs a cell whose box holds
// the point for region and position queries.
long HIPLinkedCellsAdapter::cellIndexOfPoint(const double p[3]) const {
	int i3[3];
	for (int d = 0; d < 3; ++d) {
		double x = p[d];
		if (x <= _haloBoundingBoxMin[d]) x += 0.5 * _cellLength[d];
		if (x >= _haloBoundingBoxMax[d]) x -= 0.5 * _cellLength[d];
		int i = (int)std::floor((x - _boundingBoxMin[d]) / _cellLength[d]) + _haloWidthInNumCells[d];
		i3[d] = std::min(std::max(i, 0), _cellsPerDimension[d] - 1);
	}
	return ((long)i3[2] * _cellsPerDimension[1] + i3[1]) * _cellsPerDimension[0] + i3[0];
}

// Device -> mirror: positions, velocities, orientations, angular momenta,
// forces, torques, virials and ids, each molecule in its ls1 cell.
void HIPLinkedCellsAdapter::syncToHost() {
	if (_hostValid || !_ctx || _numUploaded == 0) {
		_hostValid = true;
		return;
	}
	const int64_t n = (int64_t)_numUploaded;
	std::vector<double> r(3 * n), v(3 * n), q(4 * n), L(3 * n), F(3 * n), M(3 * n), Vi(3 * n);
	std::vector<uint64_t> ids(n);
	std::vector<int64_t> cells(n);
	check(mdx_download(_ctx, n, r.data(), v.data(), q.data(), L.data(), F.data(), M.data(), Vi.data(),
					   ids.data()),
		  "mdx_download");
	check(mdx_cell_ids(_ctx, n, cells.data()), "mdx_cell_ids");
	// the component of every molecule keeps its upload order: take it from
	// the mirror before the cells are emptied
	std::vector<Component*> comp(n, nullptr);
	{
		std::vector<std::pair<uint64_t, Component*>> byId;
		for (auto& cell : _cells)
			for (auto it = cell.iterator(); it.isValid(); ++it) byId.emplace_back(it->getID(), it->component());
		std::sort(byId.begin(), byId.end());
		for (int64_t i = 0; i < n; ++i) {
			auto f = std::lower_bound(byId.begin(), byId.end(), std::make_pair(ids[i], (Component*)nullptr));
			if (f == byId.end() || f->first != ids[i]) MARDYN_EXIT("HIPLinkedCells: molecule id lost on the device");
			comp[i] = f->second;
		}
	}
	for (auto& cell : _cells) cell.deallocateAllParticles();
	for (int64_t i = 0; i < n; ++i) {
		Molecule m(ids[i], comp[i], r[3 * i], r[3 * i + 1], r[3 * i + 2], v[3 * i], v[3 * i + 1], v[3 * i + 2],
				   q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3], L[3 * i], L[3 * i + 1], L[3 * i + 2]);
		m.setF(&F[3 * i]);
		m.setM(&M[3 * i]);
		m.setVi(&Vi[3 * i]);
		_cells.at(cells[i]).addParticle(m);
	}
	_hostValid = true;
}

// Mirror -> device, when the host may have changed it (iterators handed out,
// molecules added or deleted).
void HIPLinkedCellsAdapter::syncToDevice() {
	if (!_domainSet) MARDYN_EXIT("HIPLinkedCells: rebuild() has not been called");
	if (!_hostDirty) return;
	std::vector<double> r, v, q, L;
	std::vector<int32_t> cid;
	std::vector<uint64_t> ids;
	for (auto& cell : _cells)
		for (auto it = cell.iterator(); it.isValid(); ++it) {
			for (int d = 0; d < 3; ++d) {
				r.push_back(it->r(d));
				v.push_back(it->v(d));
				L.push_back(it->D(d));
			}
			const Quaternion& qq = it->q();
			q.insert(q.end(), {qq.qw(), qq.qx(), qq.qy(), qq.qz()});
			cid.push_back((int32_t)it->componentid());
			ids.push_back(it->getID());
		}
	_numUploaded = cid.size();
	check(mdx_upload(_ctx, (int64_t)_numUploaded, r.data(), v.data(), q.data(), L.data(), cid.data(), ids.data()),
		  "mdx_upload");
	_hostDirty = false;
}

void HIPLinkedCellsAdapter::update() {
	// LinkedCells::update + exchangeMolecules' periodic wrap on the device
	syncToDevice();
	check(mdx_update_container(_ctx), "mdx_update_container");
	deviceChanged();
}

bool HIPLinkedCellsAdapter::addParticle(Molecule& particle, bool inBoxCheckedAlready, bool checkWhetherDuplicate,
										const bool& /*rebuildCaches*/) {
	// LinkedCells::addParticle: only molecules inside the halo box
	if (!inBoxCheckedAlready) {
		for (int d = 0; d < 3; ++d)
			if (particle.r(d) < _haloBoundingBoxMin[d] || particle.r(d) >= _haloBoundingBoxMax[d]) return false;
	}
	syncToHost();
	const bool added = _cells.at(cellIndexOfPoint(particle.r_arr().data())).addParticle(particle, checkWhetherDuplicate);
	_hostDirty = _hostDirty || added;
	return added;
}

void HIPLinkedCellsAdapter::addParticles(std::vector<Molecule>& particles, bool checkWhetherDuplicate) {
	for (auto& m : particles) addParticle(m, false, checkWhetherDuplicate, false);
}

void HIPLinkedCellsAdapter::traverseCells(CellProcessor& cellProcessor) {
	auto* hip = dynamic_cast<HIPCellProcessorAdapter*>(&cellProcessor);
	if (!hip)
		MARDYN_EXIT("HIPLinkedCells: the molecules live on the device; only HIPCellProcessor can traverse them");
	cellProcessor.initTraversal();
	hip->traverseOnDevice(*this);
	cellProcessor.endTraversal();
}

void HIPLinkedCellsAdapter::traverseNonInnermostCells(CellProcessor&) {
	MARDYN_EXIT("HIPLinkedCells: overlapping communication (split traversals) is not supported");
}

void HIPLinkedCellsAdapter::traversePartialInnermostCells(CellProcessor&, unsigned int, int) {
	MARDYN_EXIT("HIPLinkedCells: overlapping communication (split traversals) is not supported");
}

ParticleIterator HIPLinkedCellsAdapter::iterator(ParticleIterator::Type t) {
	syncToHost();
	_hostDirty = true;  // the caller may change molecules through the iterator
	return ParticleIterator(t, &_cells, mardyn_get_thread_num(), mardyn_get_num_threads());
}

RegionParticleIterator HIPLinkedCellsAdapter::regionIterator(const double startRegion[3], const double endRegion[3],
															 ParticleIterator::Type type) {
	// LinkedCells::regionIterator (LinkedCells.cpp:635-682) on the mirror grid
	syncToHost();
	_hostDirty = true;
	const double* lo = type == ParticleIterator::ALL_CELLS ? _haloBoundingBoxMin : _boundingBoxMin;
	const double* hi = type == ParticleIterator::ALL_CELLS ? _haloBoundingBoxMax : _boundingBoxMax;
	double s[3], e[3];
	for (int d = 0; d < 3; ++d) {
		s[d] = std::clamp(startRegion[d], lo[d], hi[d]);
		e[d] = std::clamp(endRegion[d], lo[d], hi[d]);
	}
	const long si = cellIndexOfPoint(s), ei = cellIndexOfPoint(e);
	const int nx = _cellsPerDimension[0], ny = _cellsPerDimension[1];
	const int s3[3] = {int(si % nx), int((si / nx) % ny), int(si / ((long)nx * ny))};
	const int e3[3] = {int(ei % nx), int((ei / nx) % ny), int(ei / ((long)nx * ny))};
	const int dims[3] = {e3[0] - s3[0] + 1, e3[1] - s3[1] + 1, e3[2] - s3[2] + 1};
	return RegionParticleIterator(type, &_cells, mardyn_get_thread_num(), mardyn_get_num_threads(), (int)si, dims,
								  _cellsPerDimension, startRegion, endRegion);
}

unsigned long HIPLinkedCellsAdapter::getNumberOfParticles(ParticleIterator::Type /*t*/) {
	// no halo copies exist: every molecule is an inner one
	if (!_hostDirty) return _numUploaded;
	unsigned long n = 0;
	for (auto& cell : _cells) n += cell.getMoleculeCount();
	return n;
}

void HIPLinkedCellsAdapter::clear() {
	for (auto& cell : _cells) cell.deallocateAllParticles();
	_hostValid = true;
	_hostDirty = true;
}

void HIPLinkedCellsAdapter::deleteOuterParticles() {
	// the device keeps no halo copies; the mirror's halo cells stay empty
}

double HIPLinkedCellsAdapter::getHaloWidthForDimension(int index) const { return _cellLength[index]; }

void HIPLinkedCellsAdapter::deleteMolecule(ParticleIterator& moleculeIter, const bool& /*rebuildCaches*/) {
	moleculeIter.deleteCurrentParticle();
	_hostDirty = true;
}

double HIPLinkedCellsAdapter::getEnergy(ParticlePairsHandler*, Molecule*, CellProcessor&) {
	MARDYN_EXIT("HIPLinkedCells: test-particle energies (getEnergy) are not supported");
	return 0.0;
}

std::variant<ParticleIterator, SingleCellIterator<ParticleCell>> HIPLinkedCellsAdapter::getMoleculeAtPosition(
	const double pos[3]) {
	// LinkedCells::getMoleculeAtPosition (LinkedCells.cpp:1144-1163)
	syncToHost();
	_hostDirty = true;
	const double epsi = _cutoffRadius * 1e-6;
	auto& cell = _cells.at(cellIndexOfPoint(pos));
	for (auto it = cell.iterator(); it.isValid(); ++it)
		if (std::fabs(it->r(0) - pos[0]) <= epsi && std::fabs(it->r(1) - pos[1]) <= epsi &&
			std::fabs(it->r(2) - pos[2]) <= epsi)
			return it;
	return {};
}

unsigned long HIPLinkedCellsAdapter::initCubicGrid(std::array<unsigned long, 3> numMoleculesPerDimension,
												   std::array<double, 3> simBoxLength) {
	// the cells' own grid initialisation on the mirror, ids in cell order
	// (LinkedCells::initCubicGrid, LinkedCells.cpp:941-1002, one thread)
	syncToHost();
	unsigned long id = 0;
	for (auto& cell : _cells) {
		cell.initCubicGrid(numMoleculesPerDimension, simBoxLength);
		for (auto it = cell.iterator(); it.isValid(); ++it) it->setid(id++);
	}
	_hostDirty = true;
	return id;
}

std::string HIPLinkedCellsAdapter::getConfigurationAsString() {
	std::stringstream ss;
	ss << "{Container: HIPLinkedCells, device " << _device << ", precision "
	   << (_precision == MDX_PREC_FP64 ? "FP64" : "FP32 pair / FP64 accumulate") << "}";
	return ss.str();
}

std::vector<unsigned long> HIPLinkedCellsAdapter::getParticleCellStatistics() {
	syncToHost();
	int maxParticles = 0;
	for (auto& cell : _cells)
		if (!cell.isHaloCell()) maxParticles = std::max(maxParticles, cell.getMoleculeCount());
	return {(unsigned long)maxParticles};
}

size_t HIPLinkedCellsAdapter::getTotalSize() {
	size_t s = _cells.capacity() * sizeof(ParticleCell);
	for (auto& cell : _cells) s += cell.getMoleculeCount() * sizeof(Molecule);
	return s;
}

void HIPLinkedCellsAdapter::printSubInfo(int offset) {
	Log::global_log->info() << std::string(offset, '\t') << "host mirror cells: " << _cells.size()
							<< ", molecules on the device: " << _numUploaded << std::endl;
}

// ---------------------------------------------------- HIPCellProcessorAdapter

HIPCellProcessorAdapter::HIPCellProcessorAdapter(Domain& domain, double cutoffRadius, double LJcutoffRadius)
	: CellProcessor(cutoffRadius, LJcutoffRadius), _domain(domain) {}

void HIPCellProcessorAdapter::initTraversal() {
	for (double& m : _macro) m = 0.0;
}

void HIPCellProcessorAdapter::processCellPair(ParticleCell&, ParticleCell&, bool) {
	MARDYN_EXIT("HIPCellProcessor: cell pairs are enumerated on the device (use HIPLinkedCells)");
}

void HIPCellProcessorAdapter::processCell(ParticleCell&) {
	MARDYN_EXIT("HIPCellProcessor: cells are processed on the device (use HIPLinkedCells)");
}

double HIPCellProcessorAdapter::processSingleMolecule(Molecule*, ParticleCell&) {
	MARDYN_EXIT("HIPCellProcessor: single-molecule energies are not supported");
	return 0.0;
}

void HIPCellProcessorAdapter::traverseOnDevice(HIPLinkedCellsAdapter& container) {
	container.syncToDevice();
	const int rc = mdx_traverse(container.context(), _macro);
	if (rc != MDX_OK) MARDYN_EXIT(std::string("HIPCellProcessor: mdx_traverse: ") + mdx_last_error(container.context()));
	container.deviceChanged();
}

void HIPCellProcessorAdapter::endTraversal() {
	// VectorizedCellProcessor::endTraversal (VectorizedCellProcessor.cpp:
	// 152-154)
	_domain.setLocalVirial(_macro[MDX_VIRIAL] + 3.0 * _macro[MDX_MYRF]);
	_domain.setLocalUpot(_macro[MDX_UPOT6LJ] / 6.0 + _macro[MDX_UPOTXPOLES] + _macro[MDX_MYRF]);
}

// ------------------------------------------------------------ HIPLeapfrogAdapter

void HIPLeapfrogAdapter::readXML(XMLfileUnits& xmlconfig) {
	// Leapfrog::readXML: <timestep>
	_timestepLength = 0;
	xmlconfig.getNodeValueReduced("timestep", _timestepLength);
	Log::global_log->info() << "HIPLeapfrog timestep: " << _timestepLength << std::endl;
	if (!(_timestepLength > 0)) MARDYN_EXIT("HIPLeapfrog: timestep must be positive");
}

HIPLinkedCellsAdapter& HIPLeapfrogAdapter::device(ParticleContainer* moleculeContainer) {
	auto* c = dynamic_cast<HIPLinkedCellsAdapter*>(moleculeContainer);
	if (!c) MARDYN_EXIT("HIPLeapfrog needs the HIPLinkedCells container");
	return *c;
}

void HIPLeapfrogAdapter::eventNewTimestep(ParticleContainer* moleculeContainer, Domain*) {
	HIPLinkedCellsAdapter& c = device(moleculeContainer);
	c.syncToDevice();
	const int rc = mdx_leapfrog_pre(c.context(), _timestepLength);
	if (rc != MDX_OK) MARDYN_EXIT(std::string("HIPLeapfrog: ") + mdx_last_error(c.context()));
	c.deviceChanged();
}

void HIPLeapfrogAdapter::eventForcesCalculated(ParticleContainer* moleculeContainer, Domain* domain) {
	HIPLinkedCellsAdapter& c = device(moleculeContainer);
	c.syncToDevice();
	double sums[4] = {0, 0, 0, 0};  // sum m v^2, sum I w^2, N, rotational DOF
	const int rc = mdx_leapfrog_post(c.context(), _timestepLength, sums);
	if (rc != MDX_OK) MARDYN_EXIT(std::string("HIPLeapfrog: ") + mdx_last_error(c.context()));
	c.deviceChanged();
	// one global thermostat (id 0), as Leapfrog::transition2to3 without
	// component thermostats (Leapfrog.cpp:104-112)
	domain->setLocalSummv2(sums[0], 0);
	domain->setLocalSumIw2(sums[1], 0);
	domain->setLocalNrotDOF(0, (unsigned long)sums[2], (unsigned long)sums[3]);
}
