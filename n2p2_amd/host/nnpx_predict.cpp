// nnpx-predict: command-line equivalent of n2p2's nnp-predict
// (src/application/nnp-predict.cpp:30-262) on one GPU.  Run in a directory
// holding input.data, input.nn, scaling.data and weights.%03d.data; writes
// nnp-predict.log, energy.out, nnatoms.out, nnforces.out and output.data in
// the reference's formats (and structure.out when <info> is 1).
//
//   nnpx-predict <info> [--mixed] [--push] [--device N]
//
// --mixed selects the mixed-precision form, --push the LAMMPS (push) force
// semantics; the defaults reproduce nnp-predict (plain FP64, pull forces).
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace {

using nnpx::strpr;

int usage(const char* argv0)
{
    std::cout << "USAGE: " << argv0 << " <info> [--mixed] [--push] [--device N]\n"
              << "       <info> ... Write structure information for debugging to \"structure.out (0/1)\".\n"
              << "       Execute in directory with these NNP files present:\n"
              << "       - input.data (structure file)\n"
              << "       - input.nn (NNP settings)\n"
              << "       - scaling.data (symmetry function scaling data)\n"
              << "       - \"weights.%03d.data\" (weights files)\n";
    return 1;
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) return usage(argv[0]);
    const bool info = std::atoi(argv[1]) != 0;
    bool mixed = false, push = false;
    int device = 0;
    for (int a = 2; a < argc; ++a) {
        std::string s = argv[a];
        if (s == "--mixed") mixed = true;
        else if (s == "--push") push = true;
        else if (s == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
        else return usage(argv[0]);
    }

    std::ofstream log("nnp-predict.log");
    try {
        nnpx::Prediction prediction(device);
        prediction.setup();
        if (mixed) prediction.setPrecision(nnpx::Mode::Precision::MIXED);
        if (push) prediction.setForceMode(nnpx::Mode::ForceMode::PUSH);
        log << "\n*** PREDICTION ****************************************************************\n\n";
        log << "Reading structure file...\n";
        prediction.readStructureFromFile("input.data");
        nnpx::Structure& s = prediction.structure;
        log << strpr("Structure contains %zu atoms (%zu elements).\n", s.numAtoms, s.numElements);
        log << "Calculating NNP prediction...\n";
        prediction.predict();
        log << "\n-------------------------------------------------------------------------------\n";
        log << strpr("NNP         total energy: %16.8E\n", s.energy);
        log << "\nNNP forces:\n";
        for (auto const& a : s.atoms)
            log << strpr("%10zu %2s %16.8E %16.8E %16.8E\n", a.index + 1, prediction.elementMap[a.element].c_str(),
                         a.f[0], a.f[1], a.f[2]);
        log << "-------------------------------------------------------------------------------\n";
        log << "Writing output files...\n";

        std::ofstream f("energy.out");
        nnpx::writeFileHeader(f, {"Energy comparison."}, {10, 10, 24, 24, 24, 24},
                     {"conf", "natoms", "Eref", "Ennp", "Ediff", "E_offset"},
                     {"Configuration index (starting with 1).", "Number of atoms in configuration.",
                      "Reference potential energy.", "Potential energy predicted by NNP.",
                      "Difference in energy per atom between reference and NNP prediction.",
                      "Sum of atomic offset energies (included in column Ennp)."});
        f << strpr("%10zu %10zu %24.16E %24.16E %24.16E %24.16E\n", s.index + 1, s.numAtoms, s.energyRef, s.energy,
                   (s.energyRef - s.energy) / s.numAtoms, prediction.getEnergyOffset(s));
        f.close();

        f.open("nnatoms.out");
        nnpx::writeFileHeader(f, {"Energy contributions calculated from NNP."}, {10, 10, 3, 24, 24, 24, 24},
                     {"conf", "index", "Z", "Qref", "Qnnp", "Eref_atom", "Ennp_atom"},
                     {"Configuration index (starting with 1).", "Atom index (starting with 1).",
                      "Nuclear charge of atom.", "Reference atomic charge.", "NNP atomic charge.",
                      "Reference atomic energy contribution.",
                      "Atomic energy contribution (physical units, no mean or offset energy added)."});
        for (auto const& a : s.atoms)
            f << strpr("%10zu %10zu %3zu %24.16E %24.16E %24.16E %24.16E\n", s.index + 1, a.index + 1,
                       prediction.elementMap.atomicNumber(a.element), a.chargeRef, a.charge, 0.0, a.energy);
        f.close();

        f.open("nnforces.out");
        nnpx::writeFileHeader(f, {"Atomic force comparison (ordered by atom index)."}, {10, 10, 24, 24, 24, 24, 24, 24},
                     {"conf", "index", "fxRef", "fyRef", "fzRef", "fx", "fy", "fz"},
                     {"Configuration index (starting with 1).", "Atom index (starting with 1).",
                      "Reference force in x direction.", "Reference force in y direction.",
                      "Reference force in z direction.", "Force in x direction.", "Force in y direction.",
                      "Force in z direction."});
        for (auto const& a : s.atoms)
            f << strpr("%10zu %10zu %24.16E %24.16E %24.16E %24.16E %24.16E %24.16E\n", s.index + 1, a.index + 1,
                       a.fRef[0], a.fRef[1], a.fRef[2], a.f[0], a.f[1], a.f[2]);
        f.close();

        log << "Writing structure with NNP prediction to \"output.data\".\n";
        f.open("output.data");
        std::ofstream* fp = &f;
        s.writeToFile(fp, false);
        f.close();
        if (info) {
            f.open("structure.out");
            f << strpr("numAtoms %zu\nnumElements %zu\nisPeriodic %d\nenergy %24.16E\n", s.numAtoms, s.numElements,
                       (int)s.isPeriodic, s.energy);
            for (auto const& a : s.atoms)
                f << strpr("atom %zu element %zu numNeighbors %zu energy %24.16E\n", a.index, a.element,
                           a.numNeighbors, a.energy);
            f.close();
        }
        log << "Finished.\n";
    } catch (std::exception const& e) {
        log << e.what();
        std::cerr << e.what();
        return 1;
    }
    return 0;
}
