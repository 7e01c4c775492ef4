// mdx_run -- command-line driver of the device path, in the role of
// ls1-MarDyn's MarDyn executable for the force path: it reads an ls1 ASCII
// phase-space file, builds HIPLinkedCells + HIPCellProcessor + HIPLeapfrog
// (what <datastructure type="HIPLinkedCells"> selects), runs
// Simulation::prepare_start's initial force evaluation and then the step
// order of Simulation::simulateOneTimestep (Simulation.cpp:1076-1307) with
// the optional velocity-scaling thermostat and homogeneous long-range
// correction but no plugins, and prints the reference's per-step line with
// T, U_pot per molecule and the pressure (Simulation.cpp:1383-1386). As
// Simulation::readXML does (Simulation.cpp:514), the container is rebuilt
// right after it is built, so the cells follow LinkedCells::rebuild's float
// cell edge.
//
//   mdx_run <file.inp> --cutoff RC [--lj-cutoff RCLJ] [--dt DT] [--steps N]
//           [--precision fp64|fp32] [--device N] [--open-boundary]
//           [--checkpoint OUT.inp] [--read-write-only OUT.inp] [--nvt T] [--lrc]
//           [--rmm]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "HIPLinkedCells.h"
#include "PhaseSpaceIO.h"

using namespace ls1amd;

static void usage() {
    std::fprintf(stderr,
                 "usage: mdx_run <file.inp> --cutoff RC [--lj-cutoff RCLJ] [--dt DT] "
                 "[--steps N] [--precision fp64|fp32] [--device N] [--open-boundary] "
                 "[--checkpoint OUT.inp] [--read-write-only OUT.inp] [--nvt T] [--lrc] "
                 "[--rmm]\n");
}

int main(int argc, char **argv) {
    if (argc < 2 || !std::strcmp(argv[1], "--help") || !std::strcmp(argv[1], "-h")) {
        usage();
        return argc < 2 ? 1 : 0;
    }
    std::string inp = argv[1], checkpoint, rwonly;
    double rc = -1, rclj = -1, dt = 0.002;
    long steps = 0;
    int precision = MDX_PREC_FP64, device = 0;
    bool periodic = true, lrc = false, rmm = false;
    double Ttarget = 0.0;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--cutoff") rc = std::stod(val());
        else if (a == "--lj-cutoff") rclj = std::stod(val());
        else if (a == "--dt") dt = std::stod(val());
        else if (a == "--steps") steps = std::stol(val());
        else if (a == "--precision") precision = val() == "fp32" ? MDX_PREC_FP32_FP64ACC : MDX_PREC_FP64;
        else if (a == "--device") device = std::stoi(val());
        else if (a == "--open-boundary") periodic = false;
        else if (a == "--checkpoint") checkpoint = val();
        else if (a == "--read-write-only") rwonly = val();
        else if (a == "--nvt") Ttarget = std::stod(val());
        else if (a == "--lrc") lrc = true;
        else if (a == "--rmm") rmm = true;
        else {
            usage();
            return 1;
        }
    }
    try {
        PhaseSpace ps = readPhaseSpace(inp);
        if (!rwonly.empty()) {
            writePhaseSpace(rwonly, ps);
            return 0;
        }
        if (rc <= 0) {
            usage();
            return 1;
        }
        if (rclj <= 0) rclj = rc;
        double bmin[3] = {0, 0, 0}, bmax[3] = {ps.box[0], ps.box[1], ps.box[2]};
        HIPLinkedCells cells(bmin, bmax, rc, rclj, ps.model, periodic, device, precision);
        Domain domain(ps.box);
        HIPCellProcessor proc(domain, rc, rclj);
        HIPLeapfrog integrator(dt);
        // <thermostat type="VelocityScaling"> / <longrange type="homogeneous">
        if (Ttarget > 0 || lrc) cells.setThermostat(Ttarget, lrc);
        // Simulation::readXML rebuilds the container it has just built
        // (Simulation.cpp:514): device, precision and thermostat carry over
        cells.rebuild(bmin, bmax);
        if (cells.getPrecision() != precision || cells.getDevice() != device)
            throw std::runtime_error("rebuild lost the device or the precision");
        cells.addParticles(ps.molecules);
        domain.setglobalNumMolecules(ps.molecules.size());
        // Simulation::prepare_start (Simulation.cpp:847-999)
        cells.update();
        cells.traverseCells(proc);
        if (rmm) {  // reduced-memory build: the initial half kick (Simulation.cpp:902-937)
            if (mdx_set_rmm(cells.context(), dt) != MDX_OK)
                throw std::runtime_error(mdx_last_error(cells.context()));
        }
        {  // long-range corrections of the initial configuration
            double th[8];
            if (mdx_thermostat_state(cells.context(), th) != MDX_OK)
                throw std::runtime_error(mdx_last_error(cells.context()));
            domain.setUpotCorr(th[1]);
            domain.setVirialCorr(th[2]);
        }
        std::cout.precision(12);
        std::cout << "Simstep = 0\tU_pot = " << domain.getAverageGlobalUpot()
                  << "\tvirial = " << domain.getLocalVirial() << std::endl;
        for (long s = 1; s <= steps; ++s) {
            integrator.eventNewTimestep(&cells, &domain);
            cells.update();
            cells.traverseCells(proc);
            integrator.eventForcesCalculated(&cells, &domain);
            std::cout << "Simstep = " << s << "\tT = " << domain.getGlobalTemperature()
                      << "\tU_pot = " << domain.getAverageGlobalUpot()
                      << "\tp = " << domain.getGlobalPressure() << std::endl;
        }
        if (!checkpoint.empty()) {
            PhaseSpace out = ps;
            out.molecules = cells.getMolecules();
            out.currentTime = ps.currentTime + steps * dt;
            writePhaseSpace(checkpoint, out);
        }
    } catch (const std::exception &e) {
        std::cerr << "mdx_run: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
