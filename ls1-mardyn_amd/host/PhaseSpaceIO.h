// PhaseSpaceIO.h -- ls1 ASCII phase-space files (.inp) for the device path:
// reading mirrors ASCIIReader::readPhaseSpaceHeader / readPhaseSpace
// (src/io/ASCIIReader.cpp:52-400), writing produces the same token layout as
// Domain::writeCheckpoint (src/Domain.cpp:488-640) in the ICRVQD format, so
// a file written here reads back with either reader.
#pragma once

#include <string>
#include <vector>

#include "HIPLinkedCells.h"

namespace ls1amd {

struct PhaseSpace {
    double box[3] = {0, 0, 0};
    double temperature = 0.0;
    double currentTime = 0.0;
    ComponentModel model;
    std::vector<Molecule> molecules;
};

PhaseSpace readPhaseSpace(const std::string &path);
void writePhaseSpace(const std::string &path, const PhaseSpace &ps);

}  // namespace ls1amd
