// nnpx-dataset: command-line equivalent of n2p2's nnp-dataset
// (src/application/nnp-dataset.cpp:30-518) on one GPU.  Every structure of
// the data file is evaluated, batches of structures packed back to back in
// one device call each, instead of one structure at a time per MPI rank.
// Run in a directory holding the data file, input.nn, scaling.data and
// weights.%03d.data; writes energy.comp, forces.comp (with use_short_forces),
// output.data, sensitivity.%03d.out and nnp-dataset.log.0000 in the
// reference's formats.
//
//   nnpx-dataset <shuffle> [<data_file>] [--mixed] [--device N] [--batch-atoms M]
//
// <shuffle> only distributes structures over MPI ranks in the reference; it
// is accepted and has no effect on one GPU.  --mixed selects the
// mixed-precision form; --batch-atoms bounds the atoms per device call
// (default 100000).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace {

using nnpx::strpr;

int usage(const char* argv0)
{
    std::cout << "USAGE: " << argv0 << " <shuffle> [<data_file>] [--mixed] [--device N] [--batch-atoms M]\n"
              << "       <shuffle> ..... Randomly distribute structures to MPI processes (0/1 = no/yes;\n"
              << "                       no effect on one GPU).\n"
              << "       <data_file> ... Optional structure file name (default: input.data).\n"
              << "       Execute in directory with these NNP files present:\n"
              << "       - input.data (structure file, or <data_file> if given)\n"
              << "       - input.nn (NNP settings)\n"
              << "       - scaling.data (symmetry function scaling data)\n"
              << "       - \"weights.%03d.data\" (weights files)\n";
    return 1;
}

// Structure::updateError / Atom::updateError (Structure.cpp:1387-1415,
// Atom.cpp:347-369) and Dataset::collectError (Dataset.cpp:1712-1742).
struct Errors {
    double rmsepa = 0, rmse = 0, maepa = 0, mae = 0, frmse = 0, fmae = 0;
    size_t ne = 0, nf = 0;
    void energy(nnpx::Structure const& s)
    {
        ne++;
        double d = s.energyRef - s.energy;
        rmsepa += d * d / (double)(s.numAtoms * s.numAtoms);
        rmse += d * d;
        d = std::fabs(d);
        maepa += d / s.numAtoms;
        mae += d;
    }
    void forces(nnpx::Structure const& s)
    {
        for (auto const& a : s.atoms) {
            nf += 3;
            for (int c = 0; c < 3; ++c) {
                const double d = a.fRef[c] - a.f[c];
                frmse += d * d;
                fmae += std::fabs(d);
            }
        }
    }
    void collect()
    {
        if (ne) {
            rmsepa = std::sqrt(rmsepa / ne);
            rmse = std::sqrt(rmse / ne);
            maepa /= ne;
            mae /= ne;
        }
        if (nf) {
            frmse = std::sqrt(frmse / nf);
            fmae /= nf;
        }
    }
};

void check(nnpx_ctx* ctx, nnpx_status st)
{
    if (st != NNPX_OK) throw std::runtime_error(std::string("ERROR: ") + nnpx_last_error(ctx) + "\n");
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) return usage(argv[0]);
    std::string dataFileName = "input.data";
    bool mixed = false;
    int device = 0;
    long batchAtoms = 100000;
    for (int a = 2; a < argc; ++a) {
        std::string s = argv[a];
        if (s == "--mixed") mixed = true;
        else if (s == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
        else if (s == "--batch-atoms" && a + 1 < argc) batchAtoms = std::max(1L, std::atol(argv[++a]));
        else if (a == 2 && s.rfind("--", 0) != 0) dataFileName = s;
        else return usage(argv[0]);
    }

    std::ofstream log("nnp-dataset.log.0000");
    try {
        nnpx::Mode mode(device);
        mode.setupFromDirectory(".");
        nnpx_ctx* ctx = mode.context();
        const bool normalize = mode.useNormalization();
        const bool useForces = nnpx::settingsFileHasKeyword("input.nn", "use_short_forces");
        const double cE = mode.getConvEnergy(), cL = mode.getConvLength(), mE = mode.getMeanEnergy();
        const size_t ne = mode.getNumElements();
        int numElements = 0, maxSF = 0;
        double maxCutoff = 0.0;
        check(ctx, nnpx_model_info(ctx, &numElements, &maxSF, &maxCutoff));
        std::vector<int> nsf(ne);
        for (size_t e = 0; e < ne; ++e) check(ctx, nnpx_num_sf(ctx, (int)e, &nsf[e]));

        log << "\n*** DATA SET PREDICTION *******************************************************\n\n";
        log << (useForces ? "Energies and forces are predicted.\n" : "Only energies are predicted.\n");

        // Dataset::distributeStructures (Dataset.cpp:724-870): read, remove
        // the atomic energy offsets of the reference energy; then
        // Dataset::toNormalizedUnits.
        std::vector<nnpx::Structure> structures;
        for (auto const& lines : nnpx::readStructureBlocks(dataFileName)) {
            structures.emplace_back();
            nnpx::Structure& s = structures.back();
            s.setElementMap(mode.elementMap);
            s.index = structures.size() - 1;
            s.readFromLines(lines);
            mode.removeEnergyOffset(s);
            if (normalize) s.toNormalizedUnits(mE, cE, cL, 1.0);
        }
        log << strpr("Read %zu structures from \"%s\".\n", structures.size(), dataFileName.c_str());

        check(ctx, nnpx_set_option(ctx, NNPX_OPT_FORCE_MODE, NNPX_FORCES_PULL));
        std::vector<size_t> count(ne, 0);
        std::vector<std::vector<double>> sensMean(ne), sensMax(ne);
        for (size_t e = 0; e < ne; ++e) {
            sensMean[e].assign(nsf[e], 0.0);
            sensMax[e].assign(nsf[e], 0.0);
        }
        size_t nbatches = 0;
        for (size_t b0 = 0; b0 < structures.size();) {
            // one device call per batch of whole structures
            size_t b1 = b0;
            long atoms = 0;
            while (b1 < structures.size() && (b1 == b0 || atoms + (long)structures[b1].numAtoms <= batchAtoms))
                atoms += (long)structures[b1++].numAtoms;
            const int ns = (int)(b1 - b0);
            nnpx::PackedStructures pk;
            for (int q = 0; q < ns; ++q) pk.add(structures[b0 + q], 1.0);
            check(ctx, nnpx_set_structures(ctx, ns, pk.offset.data(), pk.position.data(), pk.element.data(),
                                           pk.box.data(), pk.periodic.data()));
            // dE/dG is always computed (the sensitivity needs it, as
            // nnp-dataset.cpp:210-215 asks for derivatives in any case)
            check(ctx, nnpx_evaluate(ctx, mixed ? NNPX_MIXED : NNPX_PLAIN_FP64, useForces ? NNPX_EVAL_FORCES : 0u));
            std::vector<double> E(ns), F(useForces ? 3 * atoms : 0), Ea(atoms), dEdG((size_t)atoms * maxSF);
            check(ctx, nnpx_get_energy(ctx, E.data()));
            check(ctx, nnpx_get_atomic_energies(ctx, Ea.data()));
            if (useForces) check(ctx, nnpx_get_forces(ctx, F.data()));
            check(ctx, nnpx_get_dEdG(ctx, dEdG.data(), maxSF));
            for (int q = 0; q < ns; ++q) {
                nnpx::Structure& s = structures[b0 + q];
                s.energy = E[q];
                for (size_t i = 0; i < s.numAtoms; ++i) {
                    nnpx::Atom& a = s.atoms[i];
                    const size_t g = (size_t)pk.offset[q] + i;
                    a.energy = Ea[g];
                    if (useForces)
                        for (int c = 0; c < 3; ++c) a.f[c] = F[3 * g + c];
                    // sensitivity (nnp-dataset.cpp:195-205)
                    const size_t e = a.element;
                    count[e]++;
                    for (int k = 0; k < nsf[e]; ++k) {
                        const double v = dEdG[g * maxSF + k];
                        sensMean[e][k] += v * v;
                        sensMax[e][k] = std::max(sensMax[e][k], std::fabs(v));
                    }
                }
            }
            b0 = b1;
            nbatches++;
        }
        log << strpr("Evaluated in %zu device call(s) (%s form).\n", nbatches, mixed ? "mixed" : "plain FP64");

        // Comparison files and errors (nnp-dataset.cpp:103-300).
        std::ofstream fe("energy.comp");
        {
            std::vector<std::string> name{"index", "N", "Eref_phys", "Ennp_phys"};
            std::vector<std::string> info{"Structure index.", "Number of atoms in structure.",
                                          "Reference potential energy (physical units, atomic energy offsets added).",
                                          "NNP potential energy (physical units, atomic energy offsets added)."};
            std::vector<size_t> size{10, 10, 16, 16};
            if (normalize) {
                name.insert(name.end(), {"Eref_int", "Ennp_int"});
                info.insert(info.end(), {"Reference potential energy (internal units).",
                                         "NNP potential energy (internal units)."});
                size.insert(size.end(), {16, 16});
            }
            nnpx::writeFileHeader(fe, {"Energy comparison."}, size, name, info);
        }
        std::ofstream ff;
        if (useForces) {
            ff.open("forces.comp");
            std::vector<std::string> name{"index_s", "index_a", "Fref_phys", "Fnnp_phys"};
            std::vector<std::string> info{"Structure index.", "Atom index (x, y, z components in consecutive lines).",
                                          "Reference force (physical units).", "NNP force (physical units)."};
            std::vector<size_t> size{10, 10, 16, 16};
            if (normalize) {
                name.insert(name.end(), {"Fref_int", "Fnnp_int"});
                info.insert(info.end(), {"Reference force (internal units).", "NNP force (internal units)."});
                size.insert(size.end(), {16, 16});
            }
            nnpx::writeFileHeader(ff, {"Force comparison."}, size, name, info);
        }
        std::ofstream fo("output.data");
        Errors err;
        const double fconv = cE / cL;
        for (auto& s : structures) {
            err.energy(s);
            fe << strpr("%10zu %10zu", s.index, s.numAtoms);
            const double off = mode.getEnergyOffset(s);
            if (normalize)
                fe << strpr(" %16.8E %16.8E %16.8E %16.8E\n", s.energyRef / cE + s.numAtoms * mE + off,
                            s.energy / cE + s.numAtoms * mE + off, s.energyRef, s.energy);
            else
                fe << strpr(" %16.8E %16.8E\n", s.energyRef + off, s.energy + off);
            if (useForces) {
                err.forces(s);
                for (auto const& a : s.atoms)
                    for (int c = 0; c < 3; ++c) {
                        ff << strpr("%10zu %10zu", s.index, a.index);
                        if (normalize) ff << strpr(" %16.8E %16.8E", a.fRef[c] / fconv, a.f[c] / fconv);
                        ff << strpr(" %16.8E %16.8E\n", a.fRef[c], a.f[c]);
                    }
            }
            if (normalize) s.toPhysicalUnits(mE, cE, cL, 1.0);
            mode.addEnergyOffset(s, false);
            std::ofstream* fp = &fo;
            s.writeToFile(fp, false);
        }
        err.collect();
        log << "Energy and force comparison in files:\n - energy.comp\n";
        if (useForces) log << " - forces.comp\n";
        log << "Predicted data set in \"output.data\"\n";
        log << "Error metrics for energies and forces:\n";
        log << std::string(118, '-') << "\n";
        log << "                      physical units                          ";
        if (normalize) log << " |                    internal units                            ";
        log << "\n";
        log << strpr("       %13s %13s %13s %13s", "RMSEpa", "RMSE", "MAEpa", "MAE");
        if (normalize) log << strpr(" | %13s %13s %13s %13s", "RMSEpa", "RMSE", "MAEpa", "MAE");
        log << "\nENERGY";
        if (normalize)
            log << strpr(" %13.5E %13.5E %13.5E %13.5E |", err.rmsepa / cE, err.rmse / cE, err.maepa / cE, err.mae / cE);
        log << strpr(" %13.5E %13.5E %13.5E %13.5E\n", err.rmsepa, err.rmse, err.maepa, err.mae);
        if (useForces) {
            log << "FORCES";
            if (normalize) log << strpr(" %13s %13.5E %13s %13.5E |", "", err.frmse / fconv, "", err.fmae / fconv);
            log << strpr(" %13s %13.5E %13s %13.5E\n", "", err.frmse, "", err.fmae);
        }
        log << std::string(118, '-') << "\n";

        // Sensitivity files (nnp-dataset.cpp:395-470).
        log << "\n*** SENSITIVITY ANALYSIS ******************************************************\n\n";
        log << "Writing sensitivity analysis data to files:\n";
        for (size_t e = 0; e < ne; ++e) {
            double sumMean = 0.0, sumMax = 0.0;
            for (int k = 0; k < nsf[e]; ++k) {
                sensMean[e][k] = std::sqrt(sensMean[e][k] / (double)count[e]);
                sumMean += sensMean[e][k];
                sumMax += sensMax[e][k];
            }
            const std::string fn = strpr("sensitivity.%03zu.out", mode.elementMap.atomicNumber(e));
            log << strpr(" - %s\n", fn.c_str());
            std::ofstream fs(fn);
            std::vector<std::string> name{"index", "sens_msa_norm", "sens_max_norm", "sens_msa_phys", "sens_max_phys"};
            std::vector<std::string> info{"Symmetry function index.",
                                          "Mean square average sensitivity (normalized, sum = 100%).",
                                          "Maximum sensitivity (normalized, sum = 100%).",
                                          "Mean square average sensitivity (physical energy units).",
                                          "Maximum sensitivity (physical energy units)."};
            std::vector<size_t> size{10, 16, 16, 16, 16};
            if (normalize) {
                name.insert(name.end(), {"sens_msa_int", "sens_max_int"});
                info.insert(info.end(), {"Mean square average sensitivity (internal units).",
                                         "Maximum sensitivity (internal units)."});
                size.insert(size.end(), {16, 16});
            }
            nnpx::writeFileHeader(fs, {strpr("Sensitivity analysis for element %2s.", mode.elementMap[e].c_str())},
                                  size, name, info);
            for (int k = 0; k < nsf[e]; ++k) {
                fs << strpr("%10d", k + 1);
                fs << strpr(" %16.8E %16.8E", sensMean[e][k] / sumMean * 100.0, sensMax[e][k] / sumMax * 100.0);
                if (normalize) fs << strpr(" %16.8E %16.8E", sensMean[e][k] / cE, sensMax[e][k] / cE);
                fs << strpr(" %16.8E %16.8E\n", sensMean[e][k], sensMax[e][k]);
            }
        }
        log << "*******************************************************************************\n";
    } catch (std::exception const& e) {
        log << e.what();
        std::cerr << e.what();
        return 1;
    }
    return 0;
}
