// nnpx-checkf: command-line equivalent of n2p2's nnp-checkf
// (src/application/nnp-checkf.cpp) on one GPU: analytic forces against
// central differences of the energy, -(E(x + delta) - E(x - delta)) /
// (2 delta), for every coordinate of every structure of input.data.  The 6N
// displaced copies of a structure (Dataset::prepareNumericForces,
// Dataset.cpp:888-950) are packed into device calls of up to --batch-atoms
// atoms instead of being spread over MPI ranks.  Always the plain FP64 form:
// the difference quotient needs energies accurate to ~1e-16 relative.
// Writes checkf-forces.out, checkf-summary.out and nnp-checkf.log.0000 in
// the reference's formats; the verdict of a structure is "WARNING!" when its
// largest error exceeds 10 delta^2 (nnp-checkf.cpp:288).
//
//   nnpx-checkf [<delta>] [--device N] [--batch-atoms M]
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace {

using nnpx::strpr;

int usage(const char* argv0)
{
    std::cout << "USAGE: " << argv0 << " <<delta>> [--device N] [--batch-atoms M]\n"
              << "       <<delta>> ... (optional) Displacement for central difference (default: 1.0e-4).\n"
              << "       Execute in directory with these NNP files present:\n"
              << "       - input.data (structure file)\n"
              << "       - input.nn (NNP settings)\n"
              << "       - scaling.data (symmetry function scaling data)\n"
              << "       - \"weights.%03d.data\" (weights files)\n";
    return 1;
}

void check(nnpx_ctx* ctx, nnpx_status st)
{
    if (st != NNPX_OK) throw std::runtime_error(std::string("ERROR: ") + nnpx_last_error(ctx) + "\n");
}

}  // namespace

int main(int argc, char* argv[])
{
    double delta = 1.0e-4;
    int device = 0;
    long batchAtoms = 1000000;
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        if (s == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
        else if (s == "--batch-atoms" && a + 1 < argc) batchAtoms = std::max(1L, std::atol(argv[++a]));
        else if (a == 1 && s.rfind("--", 0) != 0) delta = std::atof(s.c_str());
        else return usage(argv[0]);
    }

    std::ofstream log("nnp-checkf.log.0000");
    try {
        nnpx::Mode mode(device);
        mode.setupFromDirectory(".");
        nnpx_ctx* ctx = mode.context();
        // Physical <-> internal units (Structure::toNormalizedUnits /
        // toPhysicalUnits): positions scale by conv_length, energies map as
        // E_phys = E / conv_energy + N mean_energy.
        const bool normalize = mode.useNormalization();
        const double cL = normalize ? mode.getConvLength() : 1.0;
        const double cE = normalize ? mode.getConvEnergy() : 1.0;
        const double mE = normalize ? mode.getMeanEnergy() : 0.0;
        check(ctx, nnpx_set_option(ctx, NNPX_OPT_FORCE_MODE, NNPX_FORCES_PULL));

        log << "\n*** ANALYTIC/NUMERIC FORCES CHECK *********************************************\n\n";
        log << strpr("Delta for symmetric difference quotient: %11.3E\n", delta);
        std::ofstream ff("checkf-forces.out");
        log << "Individual analytic/numeric forces will be written to \"checkf-forces.out\"\n";
        nnpx::writeFileHeader(ff, {strpr("Comparison of analytic and numeric forces (delta = %11.3E).", delta)},
                              {10, 10, 3, 24, 24}, {"struct", "atom", "xyz", "F_analytic", "F_numeric"},
                              {"Structure index (starting with 1).", "Atom index (starting with 1).",
                               "Force component (0 = x, 1 = y, 2 = z).",
                               "Force computed from analytic derivative of NNP energy.",
                               "Force computed numerically from symmetric difference quotient."});
        std::ofstream fs("checkf-summary.out");
        log << "Per-structure summary of analytic/numeric force comparison will be \n"
               "written to \"checkf-summary.out\"\n";
        nnpx::writeFileHeader(
            fs, {strpr("Per-structure summary of analytic vs. numeric force comparison (delta = %11.3E).", delta)},
            {10, 16, 16, 16}, {"struct", "numForces", "meanAbsError", "maxAbsError"},
            {"Structure index (starting with 1).", "Number of forces in this structure",
             "Mean over all absolute differences between analytic and numeric forces in this structure.",
             "Maximum over all absolute differences between analytic and numeric forces in this structure."});

        auto blocks = nnpx::readStructureBlocks("input.data");
        log << strpr("Found %zu configurations in data file: %s.\n", blocks.size(), "input.data");
        log << strpr("Starting loop over %zu configurations...\n", blocks.size());
        log << "\n";
        log << strpr("                      %10s %12s %12s  verdict\n", "numForces", "meanAbsError", "maxAbsError");
        log << "-------------------------------------------------------------------\n";
        bool warning = false;
        for (size_t is = 0; is < blocks.size(); ++is) {
            nnpx::Structure original;
            original.setElementMap(mode.elementMap);
            original.readFromLines(blocks[is]);
            const size_t N = original.numAtoms;

            // analytic forces of the original (physical units)
            nnpx::PackedStructures p0;
            p0.add(original, cL);
            check(ctx, nnpx_set_structures(ctx, 1, p0.offset.data(), p0.position.data(), p0.element.data(),
                                           p0.box.data(), p0.periodic.data()));
            check(ctx, nnpx_evaluate(ctx, NNPX_PLAIN_FP64, NNPX_EVAL_FORCES));
            std::vector<double> Fa(3 * N);
            check(ctx, nnpx_get_forces(ctx, Fa.data()));
            for (double& v : Fa) v /= cE / cL;

            // energies of the 6N displaced copies: copy c displaces atom c/6,
            // component c%3, by (2 ((c%6)/3) - 1) delta (physical units)
            const size_t ncopy = 6 * N;
            std::vector<double> Ec(ncopy);
            const size_t perBatch = std::max<size_t>(1, (size_t)batchAtoms / std::max<size_t>(N, 1));
            for (size_t c0 = 0; c0 < ncopy; c0 += perBatch) {
                const size_t c1 = std::min(ncopy, c0 + perBatch);
                nnpx::PackedStructures pk;
                for (size_t c = c0; c < c1; ++c) {
                    nnpx::Structure s = original;
                    const size_t ia = c / 6, ix = c % 3;
                    const int sign = 2 * (int)((c % 6) / 3) - 1;
                    s.atoms[ia].r[ix] += sign * delta;
                    if (s.isPeriodic) s.remap(s.atoms[ia]);
                    pk.add(s, cL);
                }
                check(ctx, nnpx_set_structures(ctx, pk.count(), pk.offset.data(), pk.position.data(), pk.element.data(),
                                               pk.box.data(), pk.periodic.data()));
                check(ctx, nnpx_evaluate(ctx, NNPX_PLAIN_FP64, 0u));
                std::vector<double> E(pk.count());
                check(ctx, nnpx_get_energy(ctx, E.data()));
                for (size_t c = c0; c < c1; ++c) Ec[c] = E[c - c0] / cE + (double)N * mE;
            }

            double maxAbsError = 0.0, meanAbsError = 0.0;
            for (size_t i = 0; i < N; ++i)
                for (size_t j = 0; j < 3; ++j) {
                    // copies 6i+j (sign -1) and 6i+3+j (sign +1)
                    const double Fn = -(Ec[6 * i + 3 + j] - Ec[6 * i + j]) / (2.0 * delta);
                    const double error = std::fabs(Fn - Fa[3 * i + j]);
                    meanAbsError += error;
                    maxAbsError = std::max(error, maxAbsError);
                    ff << strpr("%10zu %10zu %3zu %24.16E %24.16E\n", is + 1, i + 1, j, Fa[3 * i + j], Fn);
                }
            const size_t numForces = 3 * N;
            meanAbsError /= numForces;
            log << strpr("Configuration %6zu: %10zu %12.3E %12.3E", is + 1, numForces, meanAbsError, maxAbsError);
            fs << strpr("%10zu %16zu %16.8E %16.8E\n", is + 1, numForces, meanAbsError, maxAbsError);
            if (maxAbsError > 10 * delta * delta) {
                warning = true;
                log << "  WARNING!\n";
            } else {
                log << "  OK.\n";
            }
        }
        log << "-------------------------------------------------------------------\n";
        if (warning)
            log << strpr("\nIMPORTANT: Some warnings were issued. By default, this happens if the\n"
                         "           maximum absolute error is higher than 10 * delta^2 (%11.3E).\n",
                         10 * delta * delta);
        log << "*******************************************************************************\n";
    } catch (std::exception const& e) {
        log << e.what();
        std::cerr << e.what();
        return 1;
    }
    return 0;
}
