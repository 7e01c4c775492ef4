// PhaseSpaceIO.cpp -- see PhaseSpaceIO.h.
#include "PhaseSpaceIO.h"

#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace ls1amd {

namespace {
// whitespace tokens, skipping lines that start with '#'
class Tokens {
public:
    explicit Tokens(std::istream &in) {
        std::string line;
        while (std::getline(in, line)) {
            size_t p = line.find_first_not_of(" \t\r");
            if (p == std::string::npos || line[p] == '#') continue;
            std::istringstream ls(line);
            std::string t;
            while (ls >> t) _t.push_back(t);
        }
    }
    const std::string &next() {
        if (_i >= _t.size()) throw std::runtime_error("phase-space file ends early");
        return _t[_i++];
    }
    double num() { return std::stod(next()); }
    long integer() { return std::stol(next()); }

private:
    std::vector<std::string> _t;
    size_t _i = 0;
};
}  // namespace

PhaseSpace readPhaseSpace(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    Tokens tk(in);
    PhaseSpace ps;
    if (tk.next() != "mardyn" || tk.next() != "trunk")
        throw std::runtime_error(path + ": not an ls1 phase-space file");
    tk.next();  // version
    long nmol = -1;
    std::string fmt = "ICRVQD";
    while (true) {
        std::string tok = tk.next();
        if (tok == "currentTime" || tok == "t") {
            ps.currentTime = tk.num();
        } else if (tok == "Temperature" || tok == "T") {
            ps.temperature = tk.num();
        } else if (tok == "Length" || tok == "L") {
            for (double &b : ps.box) b = tk.num();
        } else if (tok == "NumberOfComponents" || tok == "C") {
            int nc = (int)tk.integer();
            ps.model.ncomp = nc;
            auto &d = ps.model.desc;
            for (int i = 0; i < nc; ++i) {
                long nlj = tk.integer(), nch = tk.integer(), ndp = tk.integer(),
                     nqp = tk.integer(), nt = tk.integer();
                if (nt != 0) throw std::runtime_error("tersoff no longer supported");
                d.push_back(nlj);
                d.push_back(nch);
                d.push_back(ndp);
                d.push_back(nqp);
                for (long s = 0; s < 8 * nlj + 5 * nch + 7 * ndp + 7 * nqp + 3; ++s)
                    d.push_back(tk.num());
            }
            for (int k = 0; k < nc * (nc - 1) / 2; ++k) {
                ps.model.mix.push_back(tk.num());  // xi
                ps.model.mix.push_back(tk.num());  // eta
            }
            ps.model.epsilonRF = tk.num();
        } else if (tok == "NumberOfMolecules" || tok == "N") {
            nmol = tk.integer();
        } else if (tok == "MoleculeFormat" || tok == "M") {
            fmt = tk.next();
            break;
        } else {
            throw std::runtime_error(path + ": unexpected token '" + tok + "'");
        }
    }
    if (nmol < 0) throw std::runtime_error(path + ": NumberOfMolecules missing");
    if (fmt != "ICRVQD" && fmt != "ICRVQDV" && fmt != "ICRV" && fmt != "IRV")
        throw std::runtime_error("Unknown molecule format '" + fmt + "'");
    ps.molecules.resize((size_t)nmol);
    for (long i = 0; i < nmol; ++i) {
        Molecule &m = ps.molecules[(size_t)i];
        m.id = (uint64_t)tk.integer();
        m.cid = 0;
        if (fmt != "IRV") m.cid = (int)tk.integer() - 1;  // 1-based in the file
        if (m.cid < 0 || m.cid >= ps.model.ncomp)
            throw std::runtime_error("molecule component id out of range");
        for (double &x : m.r) x = tk.num();
        for (double &x : m.v) x = tk.num();
        if (fmt == "ICRVQD" || fmt == "ICRVQDV") {
            for (double &x : m.q) x = tk.num();
            for (double &x : m.D) x = tk.num();
        }
        if (fmt == "ICRVQDV")
            for (int k = 0; k < 3; ++k) tk.num();
    }
    return ps;
}

void writePhaseSpace(const std::string &path, const PhaseSpace &ps) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write " + path);
    out << "mardyn trunk 20100525\n";
    out << std::setprecision(17);
    out << "currentTime\t" << ps.currentTime << "\n";
    out << " Length\t" << ps.box[0] << " " << ps.box[1] << " " << ps.box[2] << "\n";
    out << " Temperature\t" << ps.temperature << "\n";
    out << " NumberOfComponents\t" << ps.model.ncomp << "\n";
    const auto &d = ps.model.desc;
    size_t p = 0;
    for (int c = 0; c < ps.model.ncomp; ++c) {
        long n[4] = {(long)d[p], (long)d[p + 1], (long)d[p + 2], (long)d[p + 3]};
        out << n[0] << "\t" << n[1] << "\t" << n[2] << "\t" << n[3] << "\t0\n";
        p += 4;
        const int width[4] = {8, 5, 7, 7};
        for (int t = 0; t < 4; ++t)
            for (long s = 0; s < n[t]; ++s) {
                for (int k = 0; k < width[t]; ++k) out << (k ? " " : "") << d[p + k];
                out << "\n";
                p += width[t];
            }
        out << d[p] << " " << d[p + 1] << " " << d[p + 2] << "\n";
        p += 3;
    }
    for (size_t k = 0; k + 1 < ps.model.mix.size(); k += 2)
        out << ps.model.mix[k] << " " << ps.model.mix[k + 1] << "\n";
    out << ps.model.epsilonRF << "\n";
    out << " NumberOfMolecules\t" << ps.molecules.size() << "\n";
    out << " MoleculeFormat\tICRVQD\n";
    for (const Molecule &m : ps.molecules) {
        out << m.id << "\t" << (m.cid + 1);
        for (double x : m.r) out << "\t" << x;
        for (double x : m.v) out << "\t" << x;
        for (double x : m.q) out << "\t" << x;
        for (double x : m.D) out << "\t" << x;
        out << "\n";
    }
}

}  // namespace ls1amd
